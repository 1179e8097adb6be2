// Drives the C++ facade (include/cgspline/cgspline.hpp) the way the
// reference's own callers use rt::BVH, rt::PerfectHashFunction, rt::Scene and
// rt::RayTracer (MainWindow::convertScene / renderScene).  Inputs and outputs
// are little-endian binary files written and checked by tests/test_facade.py.
//
//   facade_check bvh    <in> <out>   BVH::build on the host (no device)
//   facade_check hash   <in> <out>   PerfectHashFunction + BVH::buildHashTable
//   facade_check render <in> <out>   Scene::buildBVH + RayTracer::render
//   facade_check load_be <in> <out>  Surface::load_be on a `.be` model
//   facade_check load_se <in> <out>  Surface::load_se on a `.se` surface
//   facade_check errors              error behaviour without a device
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

#include "cgspline/cgspline.hpp"

using namespace cg;

namespace
{

struct Reader
{
    std::ifstream in;
    explicit Reader(const char* p) : in(p, std::ios::binary)
    {
        if (!in) throw std::runtime_error(std::string("cannot open ") + p);
    }
    template<class T> T get()
    {
        T v;
        in.read((char*)&v, sizeof v);
        if (!in) throw std::runtime_error("short input");
        return v;
    }
    template<class T> std::vector<T> vec(size_t n)
    {
        std::vector<T> v(n);
        in.read((char*)v.data(), sizeof(T) * n);
        if (!in) throw std::runtime_error("short input");
        return v;
    }
};

struct Writer
{
    std::ofstream out;
    explicit Writer(const char* p) : out(p, std::ios::binary) {}
    template<class T> void put(const T& v) { out.write((const char*)&v, sizeof v); }
    template<class T> void vec(std::span<const T> v) { out.write((const char*)v.data(), sizeof(T) * v.size()); }
};

int runBVH(const char* inPath, const char* outPath)
{
    Reader r(inPath);
    const uint32_t n = r.get<uint32_t>(), epn = r.get<uint32_t>();
    auto el = r.vec<rt::BVH::ElementData>(n);
    rt::BVH bvh;
    bvh.build(el, epn);
    Writer w(outPath);
    w.put((uint32_t)bvh.nodes().size());
    w.put((uint32_t)bvh.indices().size());
    w.vec(bvh.nodes());
    w.vec(bvh.keys());
    w.vec(bvh.indices());
    return 0;
}

int runHash(const char* inPath, const char* outPath)
{
    Reader r(inPath);
    const uint32_t n = r.get<uint32_t>(), epn = r.get<uint32_t>();
    auto el = r.vec<rt::BVH::ElementData>(n);
    rt::Device dev(0);
    rt::BVH bvh(&dev);
    bvh.build(el, epn);
    bvh.buildHashTable();
    const rt::PerfectHashFunction& h = *bvh.hash();
    // the table maps every key's hash to the key's node, EMPTY elsewhere
    auto keys = bvh.keys();
    std::vector<uint32_t> hashes(keys.size());
    for (size_t i = 0; i < keys.size(); i++)
    {
        hashes[i] = h.hash(keys[i]);
        if (bvh.table()[hashes[i]] != i)
        {
            std::fprintf(stderr, "table[h(key %zu)] = %u\n", i, bvh.table()[hashes[i]]);
            return 3;
        }
    }
    // a second function over the same keys from the host span
    rt::PerfectHashFunction again(dev);
    if (again.build(keys) != 0 || again.tableSize() != h.tableSize()) return 4;
    Writer w(outPath);
    w.put(h.parameters());
    w.vec(std::span<const uint32_t>(hashes));
    w.vec(bvh.table());
    return 0;
}

int runRender(const char* inPath, const char* outPath)
{
    Reader r(inPath);
    rt::Scene scene;
    const uint32_t nsurf = r.get<uint32_t>(), nobj = r.get<uint32_t>(), nlight = r.get<uint32_t>();
    std::vector<std::pair<std::vector<float>, std::vector<uint32_t>>> surfaces;
    for (uint32_t s = 0; s < nsurf; s++)
    {
        const uint32_t nv = r.get<uint32_t>(), np = r.get<uint32_t>();
        auto v = r.vec<float>(4 * (size_t)nv);
        auto ix = r.vec<uint32_t>(16 * (size_t)np);
        scene.addSurface(v, ix);
        surfaces.emplace_back(std::move(v), std::move(ix));
    }
    for (uint32_t o = 0; o < nobj; o++)
    {
        rt::Scene::Object ob {};
        ob.surface = r.get<uint32_t>();
        auto l2w = r.vec<float>(16), w2l = r.vec<float>(16), m = r.vec<float>(8);
        std::memcpy(ob.local2world, l2w.data(), 64);
        std::memcpy(ob.world2local, w2l.data(), 64);
        std::memcpy(ob.material.diffuse, &m[0], 12);
        std::memcpy(ob.material.specular, &m[3], 12);
        ob.material.metalness = m[6];
        ob.material.roughness = m[7];
        scene.objects.push_back(ob);
    }
    for (uint32_t l = 0; l < nlight; l++)
    {
        auto v = r.vec<float>(12);
        rt::Light li {};
        std::memcpy(&li, v.data(), sizeof li);
        scene.lights.push_back(li);
    }
    auto bg = r.vec<float>(3), amb = r.vec<float>(3);
    std::memcpy(scene.backgroundColor, bg.data(), 12);
    std::memcpy(scene.ambientLight, amb.data(), 12);
    scene.hashSeed = r.get<uint32_t>();
    rt::Camera cam {};
    auto tw = r.vec<float>(16), pos = r.vec<float>(3);
    std::memcpy(cam.toWorld, tw.data(), 64);
    std::memcpy(cam.position, pos.data(), 12);
    cam.fieldOfView = r.get<float>();
    const uint32_t width = r.get<uint32_t>(), height = r.get<uint32_t>();
    const float eps = r.get<float>();
    const int32_t depth = r.get<int32_t>();

    rt::RayTracer tracer({.eps = eps, .flipYAxis = false, .recursionDepth = depth, .device = 0});
    const rt::Device& dev = tracer.device();
    scene.buildBVH(dev);
    rt::DeviceArray<uint32_t> pixels(dev, (size_t)width * height);
    rt::Frame frame {pixels.data(), width, height};
    tracer.render(&frame, &cam, &scene);
    // a second render while the first may still run is refused with a warning
    tracer.render(&frame, &cam, &scene);
    dev.synchronize();
    if (tracer.running()) return 5;
    auto host = pixels.toHost();
    Writer w(outPath);
    w.vec(std::span<const uint32_t>(host));

    // The reference's usage across instances: a scene built on its own
    // (the GPU's shared Device) and rendered by two RayTracers at once, each
    // on a stream of its own.
    rt::Scene shared;
    shared.objects = scene.objects;
    shared.lights = scene.lights;
    std::memcpy(shared.backgroundColor, scene.backgroundColor, 12);
    std::memcpy(shared.ambientLight, scene.ambientLight, 12);
    shared.hashSeed = scene.hashSeed;
    for (uint32_t s = 0; s < nsurf; s++) shared.addSurface(surfaces[s].first, surfaces[s].second);
    shared.buildBVH();
    rt::RayTracer t1({.eps = eps, .flipYAxis = false, .recursionDepth = depth, .device = 0});
    rt::RayTracer t2({.eps = eps, .flipYAxis = false, .recursionDepth = depth, .device = 0});
    rt::DeviceArray<uint32_t> p1(t1.device(), (size_t)width * height), p2(t2.device(), (size_t)width * height);
    rt::Frame f1 {p1.data(), width, height}, f2 {p2.data(), width, height};
    rt::stream_t s1 = t1.device().createStream(), s2 = t2.device().createStream();
    for (int k = 0; k < 3; k++)
    {
        t1.render(&f1, &cam, &shared, s1);
        t2.render(&f2, &cam, &shared, s2);
        t1.device().synchronize(s1);
        t2.device().synchronize(s2);
    }
    t1.device().destroyStream(s1);
    t2.device().destroyStream(s2);
    w.vec(std::span<const uint32_t>(p1.toHost()));
    w.vec(std::span<const uint32_t>(p2.toHost()));

    // The Monte Carlo render (the reference's USE_MONTECARLO_SAMPLING build)
    // with two samples per pixel and the default seed.
    rt::RayTracer mc({.samples = 2, .eps = eps, .flipYAxis = false, .recursionDepth = depth, .device = 0,
                      .monteCarlo = true});
    rt::DeviceArray<uint32_t> p3(mc.device(), (size_t)width * height);
    rt::Frame f3 {p3.data(), width, height};
    mc.render(&f3, &cam, &shared);
    mc.device().synchronize();
    w.vec(std::span<const uint32_t>(p3.toHost()));
    return 0;
}

int runLoadBe(const char* inPath, const char* outPath)
{
    std::ifstream in(inPath, std::ios::binary);
    if (!in) throw std::runtime_error(std::string("cannot open ") + inPath);
    const std::shared_ptr<Surface> s = Surface::load_be(in);
    Writer w(outPath);
    w.put((uint32_t)(s->points().size() / 4));
    w.put(s->count());
    w.vec(s->points());
    w.vec(s->indices());
    return 0;
}

// Surface::load_se on a `.se` surface: counts, points, the 16 leaf indices
// per patch and the PatchBVH element boxes.
int runLoadSe(const char* inPath, const char* outPath)
{
    std::ifstream in(inPath, std::ios::binary);
    if (!in) throw std::runtime_error(std::string("cannot open ") + inPath);
    const std::shared_ptr<Surface> s = Surface::load_se(in);
    Writer w(outPath);
    w.put((uint32_t)(s->points().size() / 4));
    w.put(s->count());
    w.vec(s->points());
    w.vec(s->indices());
    const auto el = s->elements();
    w.vec(std::span<const float>((const float*)el.data(), el.size() * 10));
    return 0;
}

// The per-call functions and PatchBVH the way a picking caller uses them:
// a .be model, rays (n × 8 floats: origin, tMax, direction, pad); for each
// ray: PatchBVH::intersect(ray, hit) (found, distance, u, v, patch),
// PatchBVH::intersect(ray), spline::doBezierClipping and doSubdivision of the
// ray against patch k mod count (found, distance, u, v, each), and
// spline::evaluate(CGS_EVAL_NORMAL) of that patch at (0.3, 0.6).
int runPick(const char* modelPath, const char* raysPath, const char* outPath)
{
    std::ifstream in(modelPath, std::ios::binary);
    if (!in) throw std::runtime_error(std::string("cannot open ") + modelPath);
    const std::shared_ptr<Surface> s = Surface::load_be(in);
    std::ifstream rf(raysPath, std::ios::binary | std::ios::ate);
    if (!rf) throw std::runtime_error(std::string("cannot open ") + raysPath);
    const size_t nbytes = (size_t)rf.tellg();
    rf.seekg(0);
    std::vector<float> r(nbytes / 4);
    rf.read((char*)r.data(), (std::streamsize)(r.size() * 4));
    const PatchBVH bvh(s.get());
    Writer w(outPath);
    const uint32_t n = (uint32_t)(r.size() / 8);
    for (uint32_t k = 0; k < n; k++)
    {
        const float* q = &r[8 * k];
        Ray3f ray {{q[0], q[1], q[2]}, {q[4], q[5], q[6]}, q[3]};
        Intersection hit;
        const bool f = bvh.intersect(ray, hit);
        w.put((uint32_t)f);
        w.put(hit.distance);
        w.put(hit.p[0]);
        w.put(hit.p[1]);
        w.put(hit.triangleIndex);
        if (f && hit.object != &bvh) return 20;
        w.put((uint32_t)bvh.intersect(ray));
        const uint32_t* patch = s->indices().data() + 16 * (size_t)(k % s->count());
        for (int m = 0; m < 2; m++)
        {
            Intersection h;
            h.distance = q[3];
            const bool g = m == 0 ? spline::doBezierClipping(h, ray, s->points().data(), patch)
                                  : spline::doSubdivision(h, ray, s->points().data(), patch);
            w.put((uint32_t)g);
            w.put(h.distance);
            w.put(h.p[0]);
            w.put(h.p[1]);
        }
        const auto nrm = spline::evaluate(CGS_EVAL_NORMAL, s->points().data(), patch, 0.3f, 0.6f);
        for (float x : nrm) w.put(x);
    }
    return 0;
}

int runErrors()
{
    // Surface::load_be on a truncated model: runtime_error with the
    // reference's message
    try
    {
        std::istringstream bad("3\n1,2");
        Surface::load_be(bad);
        return 10;
    }
    catch (const std::runtime_error& e)
    {
        if (std::string(e.what()) != "Parsing patch failed") return 11;
    }
    // BVH::build over nothing: logic_error("invalid") like BVH::wrap
    try
    {
        rt::BVH bvh;
        std::vector<rt::BVH::ElementData> none;
        bvh.build(none);
        return 6;
    }
    catch (const std::logic_error& e)
    {
        if (std::string(e.what()) != "invalid") return 7;
    }
    // no device here: opening one throws rt::Error with CGS_ERR_NO_DEVICE
    try
    {
        rt::Device dev(0);
        return 8;
    }
    catch (const rt::Error& e)
    {
        if (e.status != CGS_ERR_NO_DEVICE) return 9;
    }
    return 0;
}

} // namespace

int main(int argc, char** argv)
{
    try
    {
        const std::string mode = argc > 1 ? argv[1] : "";
        if (mode == "bvh" && argc == 4) return runBVH(argv[2], argv[3]);
        if (mode == "hash" && argc == 4) return runHash(argv[2], argv[3]);
        if (mode == "render" && argc == 4) return runRender(argv[2], argv[3]);
        if (mode == "load_be" && argc == 4) return runLoadBe(argv[2], argv[3]);
        if (mode == "load_se" && argc == 4) return runLoadSe(argv[2], argv[3]);
        if (mode == "errors") return runErrors();
        if (mode == "pick" && argc == 5) return runPick(argv[2], argv[3], argv[4]);
        std::fprintf(stderr, "usage: facade_check bvh|hash|render|load_be <in> <out> | pick <model> <rays> <out> | errors\n");
        return 2;
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "facade_check: %s\n", e.what());
        return 1;
    }
}
