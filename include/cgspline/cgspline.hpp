// SPDX-License-Identifier: MIT
//
// cgspline.hpp — C++ facade of libcgs with the reference's source-level names
// (namespaces cg::spline and cg::rt), header-only, over the C ABI in cgs.h.
//
// The facade keeps the reference's argument meaning and error behaviour:
//   * PerfectHashFunction::build returns non-zero on failure
//     (rt/PerfectHashFunction.h:47-53);
//   * BVH::buildHashTable throws std::runtime_error when the hash fails
//     (rt/BVH.cpp:24-38); BVH::build throws std::logic_error("invalid") on an
//     empty split (rt/BVH.cpp:84-88);
//   * RayTracer::render on a busy tracer prints a warning and does nothing
//     (rt/RayTracer.cu:109-115);
//   * every other failure of the library throws cg::rt::Error (the reference
//     aborts on CUDA errors, CUDAUtility.cpp:8-22).
// Device memory is plain HIP device memory; streams are hipStream_t.
//
// What the C ABI cannot carry stays out: hashIntersect/intersect with a
// user predicate (templates over arbitrary callables) are instantiated inside
// the library for the Bézier-surface predicate (trace and render entry
// points) and are not exposed as templates.
//
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <istream>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../cgs.h"

struct ihipStream_t;

namespace cg
{

// cg::Surface (Surface.h:130-164) for one group of bicubic Bézier patches:
// points (x, y, z, w) and 16 indices per patch, point(i, j) = indices[4j+i].
// Ref<Surface> is a shared pointer here.
class Surface
{
public:
    Surface() = default;
    Surface(std::vector<float> points4, std::vector<uint32_t> indices)
        : _points(std::move(points4)), _indices(std::move(indices))
    {}

    uint32_t count() const { return (uint32_t)(_indices.size() / 16); }
    std::span<const float> points() const { return _points; }   // 4 floats per point
    std::span<const uint32_t> indices() const { return _indices; }

    // Surface::load_be (Surface.cpp:149-213): reads a `.be` model from the
    // stream; malformed text throws std::runtime_error with the reference's
    // messages (an index equal to the point count included, which the
    // reference lets through).
    static std::shared_ptr<Surface> load_be(std::istream& input)
    {
        const std::string text {std::istreambuf_iterator<char>(input), std::istreambuf_iterator<char>()};
        float* v = nullptr;
        uint32_t* ix = nullptr;
        uint32_t nv = 0, np = 0;
        if (cgs_load_be(text.data(), text.size(), &v, &nv, &ix, &np) != CGS_OK)
            throw std::runtime_error(cgs_last_error());
        auto s = std::make_shared<Surface>(std::vector<float>(v, v + 4 * (size_t)nv),
                                           std::vector<uint32_t>(ix, ix + 16 * (size_t)np));
        cgs_free_host(v);
        cgs_free_host(ix);
        return s;
    }

    // Surface::load_se (Surface.cpp:214-348): reads a `.se` element surface.
    // Its patches keep the reference's table order; indices() holds the 16
    // indices PatchBVH's leaf test reads for each (cgs_patch_table_patches)
    // and elements() the boxes PatchBVH gives them (cgs_patch_table_elements).
    // Malformed text throws std::runtime_error with the reference's messages.
    static std::shared_ptr<Surface> load_se(std::istream& input)
    {
        const std::string text {std::istreambuf_iterator<char>(input), std::istreambuf_iterator<char>()};
        cgs_patch_table* t = nullptr;
        if (cgs_load_se(text.data(), text.size(), &t) != CGS_OK) throw std::runtime_error(cgs_last_error());
        std::vector<uint32_t> idx(16 * (size_t)t->npatches);
        std::vector<cgs_bvh_element> el(t->npatches);
        const bool ok = cgs_patch_table_patches(t, idx.data()) == CGS_OK && cgs_patch_table_elements(t, el.data()) == CGS_OK;
        auto s = std::make_shared<Surface>(std::vector<float>(t->points, t->points + 4 * (size_t)t->npoints), std::move(idx));
        cgs_patch_table_free(t);
        if (!ok) throw std::runtime_error(cgs_last_error());
        s->_elements = std::move(el);
        return s;
    }

    // The PatchBVH element boxes of a `.se` surface; empty for Bézier patches,
    // which PatchBVH boxes by their control points.
    std::span<const cgs_bvh_element> elements() const { return _elements; }

private:
    std::vector<float> _points;
    std::vector<uint32_t> _indices;
    std::vector<cgs_bvh_element> _elements;
};

namespace rt
{

using stream_t = ::ihipStream_t*;   // hipStream_t

class Error : public std::runtime_error
{
public:
    Error(cgs_status s, const std::string& what)
        : std::runtime_error(what + ": " + cgs_status_string(s) + " (" + cgs_last_error() + ")"), status(s)
    {}
    cgs_status status;
};

inline void check(cgs_status s, const char* what)
{
    if (s != CGS_OK) throw Error(s, what);
}

// One opened GPU (RayTracer::Options::device).  Owns the library context:
// work queues, region stacks and staging buffers of the device path.  Data
// made resident through one Device (scenes) may be used through any other
// Device of the same GPU.
class Device
{
public:
    explicit Device(int device = 0) { check(cgs_open(device, &_ctx), "cgs_open"); }

    // A Device per GPU for callers that do not bring their own (scene
    // builds, PatchBVH): opened on first use and kept until the process ends.
    static const Device& shared(int device = 0)
    {
        static std::mutex mu;
        static std::map<int, Device*> open;
        std::lock_guard lk(mu);
        Device*& d = open[device];
        if (!d) d = new Device(device);
        return *d;
    }
    ~Device()
    {
        if (_ctx) (void)cgs_close(_ctx);
    }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    Device(Device&& o) noexcept : _ctx(std::exchange(o._ctx, nullptr)) {}

    cgs_ctx* handle() const noexcept { return _ctx; }

    void* allocate(size_t bytes) const
    {
        void* p = nullptr;
        check(cgs_malloc(_ctx, bytes, &p), "cgs_malloc");
        return p;
    }
    void deallocate(void* p) const noexcept { (void)cgs_free(_ctx, p); }
    void upload(void* dst, const void* src, size_t bytes, stream_t s = nullptr) const
    {
        check(cgs_memcpy_h2d(_ctx, dst, src, bytes, s), "cgs_memcpy_h2d");
    }
    void download(void* dst, const void* src, size_t bytes, stream_t s = nullptr) const
    {
        check(cgs_memcpy_d2h(_ctx, dst, src, bytes, s), "cgs_memcpy_d2h");
    }
    void synchronize(stream_t s = nullptr) const { check(cgs_stream_sync(_ctx, s), "cgs_stream_sync"); }
    stream_t createStream() const
    {
        void* s = nullptr;
        check(cgs_stream_create(_ctx, &s), "cgs_stream_create");
        return (stream_t)s;
    }
    void destroyStream(stream_t s) const { check(cgs_stream_destroy(_ctx, s), "cgs_stream_destroy"); }

private:
    cgs_ctx* _ctx = nullptr;
};

// Device array of T (the reference's managed Array/Buffer, here explicit
// device memory owned by one Device).
template<class T>
class DeviceArray
{
public:
    DeviceArray() = default;
    DeviceArray(const Device& d, size_t n) : _dev(&d), _n(n) { _p = (T*)d.allocate(sizeof(T) * (n ? n : 1)); }
    DeviceArray(const Device& d, std::span<const T> host, stream_t s = nullptr) : DeviceArray(d, host.size())
    {
        if (_n) d.upload(_p, host.data(), sizeof(T) * _n, s);
    }
    ~DeviceArray() { reset(); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    DeviceArray(DeviceArray&& o) noexcept
        : _dev(o._dev), _p(std::exchange(o._p, nullptr)), _n(std::exchange(o._n, 0))
    {}
    DeviceArray& operator=(DeviceArray&& o) noexcept
    {
        if (this != &o)
        {
            reset();
            _dev = o._dev;
            _p = std::exchange(o._p, nullptr);
            _n = std::exchange(o._n, 0);
        }
        return *this;
    }
    void reset() noexcept
    {
        if (_p) _dev->deallocate(_p);
        _p = nullptr;
        _n = 0;
    }
    T* data() const noexcept { return _p; }
    size_t size() const noexcept { return _n; }
    std::vector<T> toHost(stream_t s = nullptr) const
    {
        std::vector<T> h(_n);
        if (_n) _dev->download(h.data(), _p, sizeof(T) * _n, s);
        if (s) _dev->synchronize(s);
        return h;
    }

private:
    const Device* _dev = nullptr;
    T* _p = nullptr;
    size_t _n = 0;
};

// ---------------------------------------------------------------------------
// rt::PerfectHashFunction (rt/PerfectHashFunction.h:30-70)
// ---------------------------------------------------------------------------
class PerfectHashFunction
{
public:
    static constexpr auto EMPTY = uint32_t(-1);

    explicit PerfectHashFunction(const Device& device, uint32_t seed = 42) : _dev(&device), _seed(seed) {}

    // Builds h(k) = (A·k + B + D[k & mask]) mod T for the keys (host span, as
    // the reference's managed span is host-readable).  0 on success, non-zero
    // on failure like the reference (PerfectHashFunction.cu:486-503).
    int build(std::span<const uint32_t> keys, stream_t stream = nullptr)
    {
        DeviceArray<uint32_t> k(*_dev, keys, stream);
        return buildDevice(k.data(), (uint32_t)keys.size(), stream);
    }

    // Same, keys already in device memory.
    int buildDevice(const uint32_t* deviceKeys, uint32_t n, stream_t stream = nullptr)
    {
        _displacement = DeviceArray<uint32_t>(*_dev, cgs_phf_displacement_size(n));
        cgs_status s = cgs_phf_build(_dev->handle(), deviceKeys, n, _seed, _displacement.data(), &_p, stream);
        if (s == CGS_ERR_HASH_FAILED || s == CGS_ERR_HASH_DUPLICATE) return 1;
        check(s, "PerfectHashFunction::build");
        _host = _displacement.toHost(stream);
        _dev->synchronize(stream);
        return 0;
    }

    uint32_t operator()(uint32_t key) const { return hash(key); }
    // Host evaluation (the reference's hash is __host__ __device__).
    uint32_t hash(uint32_t key) const { return (_p.A * key + _p.B + _host[key & _p.mask]) % _p.table_size; }

    size_t displacementSize() const noexcept { return _p.displacement_size; }
    uint32_t tableSize() const noexcept { return _p.table_size; }
    const cgs_phf& parameters() const noexcept { return _p; }
    const uint32_t* deviceDisplacement() const noexcept { return _displacement.data(); }

private:
    const Device* _dev;
    uint32_t _seed;
    cgs_phf _p {31, 19, 0, 0, 0, 0};
    DeviceArray<uint32_t> _displacement;
    std::vector<uint32_t> _host;
};

// ---------------------------------------------------------------------------
// rt::BVH (rt/BVH.h:20-80)
// ---------------------------------------------------------------------------
class BVH
{
public:
    using key_t = uint32_t;
    using Node = cgs_bvh_node;
    using ElementData = cgs_bvh_element;
    static constexpr uint32_t EMPTY = CGS_EMPTY;

    explicit BVH(const Device* device = nullptr) : _dev(device) {}

    constexpr auto root() const { return 0; }
    std::span<const Node> nodes() const noexcept { return _nodes; }
    std::span<const uint32_t> indices() const noexcept { return _indices; }
    std::span<const key_t> keys() const noexcept { return _keys; }
    std::span<const uint32_t> table() const noexcept { return _table; }

    // Median split on the host (rt/BVH.cpp:9-108).  Unlike the reference the
    // element span is not reordered in place; the result is the same.
    void build(std::span<ElementData> elements, uint32_t elementsPerNode = 1)
    {
        Node* n = nullptr;
        uint32_t* k = nullptr;
        uint32_t* idx = nullptr;
        uint32_t nn = 0, ni = 0;
        cgs_status s = cgs_bvh_build(elements.data(), (uint32_t)elements.size(), elementsPerNode, &n, &k, &nn, &idx, &ni);
        if (s == CGS_ERR_EMPTY_SPLIT) throw std::logic_error("invalid");
        check(s, "BVH::build");
        _nodes.assign(n, n + nn);
        _keys.assign(k, k + nn);
        _indices.assign(idx, idx + ni);
        cgs_free_host(n);
        cgs_free_host(k);
        cgs_free_host(idx);
        _table.clear();
    }

    // Perfect hash over the node keys and table[h(key_i)] = i
    // (rt/BVH.cpp:24-38), built on the device.
    void buildHashTable(stream_t stream = nullptr)
    {
        if (!_dev) throw std::logic_error("BVH::buildHashTable: no device");
        _hash.emplace(*_dev);
        DeviceArray<uint32_t> k(*_dev, std::span<const uint32_t>(_keys), stream);
        if (_hash->buildDevice(k.data(), (uint32_t)_keys.size(), stream) != 0)
            throw std::runtime_error("could not build the hash table");
        DeviceArray<uint32_t> t(*_dev, _hash->tableSize());
        check(cgs_phf_fill_table(_dev->handle(), &_hash->parameters(), _hash->deviceDisplacement(), k.data(),
                                 (uint32_t)_keys.size(), t.data(), stream),
              "BVH::buildHashTable");
        _table = t.toHost(stream);
        _dev->synchronize(stream);
    }

    const PerfectHashFunction* hash() const noexcept { return _hash ? &*_hash : nullptr; }

private:
    const Device* _dev;
    std::vector<Node> _nodes;
    std::vector<key_t> _keys;
    std::vector<uint32_t> _indices;
    std::vector<uint32_t> _table;
    std::optional<PerfectHashFunction> _hash;
};

// ---------------------------------------------------------------------------
// rt::Scene (rt/Scene.h:16-159): objects, lights, Bézier surfaces
// ---------------------------------------------------------------------------
struct Material
{
    float diffuse[3];
    float specular[3];
    float metalness;
    float roughness;
};

struct Light
{
    float color[3];
    float strength;
    float position[3];
    float range = 0;
    float direction[3];
    float angle = 1.57079637f;   // > 1.57: point light
};

struct Camera
{
    float toWorld[16];   // camera→world, column-major
    float position[3];
    float fieldOfView;   // vertical, radians
};

// Scene content as MainWindow::convertScene fills rt::Scene
// (MainWindow.cpp:363-492); buildBVH() builds every BVH and hash and makes
// the scene resident on the device (Scene::buildBVH, rt/Scene.cu:39-53).
class Scene
{
public:
    struct Object
    {
        uint32_t surface;   // CGS_EMPTY: a mesh (no geometry here; empty bounds, never hit, as in the reference)
        float local2world[16];
        float world2local[16];
        Material material;
    };

    float backgroundColor[3] {0.1f, 0.1f, 0.1f};
    float ambientLight[3] {0.1f, 0.1f, 0.1f};
    uint32_t hashSeed = 42;
    std::vector<Object> objects;
    std::vector<Light> lights;

    // A BezierSurface (rt/Primitive.h:90-116): vertices (x,y,z,w) and 16
    // indices per patch, point(i,j) = indices[4j+i].  Returns its index.
    uint32_t addSurface(std::span<const float> vertices4, std::span<const uint32_t> indices)
    {
        _vbase.push_back(_vbase.back() + (uint32_t)(vertices4.size() / 4));
        _pbase.push_back(_pbase.back() + (uint32_t)(indices.size() / 16));
        _vertices.insert(_vertices.end(), vertices4.begin(), vertices4.end());
        _indices.insert(_indices.end(), indices.begin(), indices.end());
        return (uint32_t)_vbase.size() - 2;
    }
    uint32_t addSurface(const cg::Surface& surface) { return addSurface(surface.points(), surface.indices()); }

    // Builds on the GPU's shared Device (Device::shared); any RayTracer on
    // that GPU can render the result.
    void buildBVH(int device = 0) { buildBVH(Device::shared(device)); }

    void buildBVH(const Device& device)
    {
        std::vector<uint32_t> surf;
        std::vector<float> l2w, w2l, mat, lt;
        for (const Object& o : objects)
        {
            surf.push_back(o.surface);
            l2w.insert(l2w.end(), o.local2world, o.local2world + 16);
            w2l.insert(w2l.end(), o.world2local, o.world2local + 16);
            const Material& m = o.material;
            mat.insert(mat.end(), {m.diffuse[0], m.diffuse[1], m.diffuse[2], m.specular[0], m.specular[1],
                                   m.specular[2], m.metalness, m.roughness});
        }
        for (const Light& l : lights)
            lt.insert(lt.end(), {l.color[0], l.color[1], l.color[2], l.strength, l.position[0], l.position[1],
                                 l.position[2], l.range, l.direction[0], l.direction[1], l.direction[2], l.angle});
        cgs_scene_desc d {};
        d.vertices = _vertices.data();
        d.indices = _indices.data();
        d.surf_vbase = _vbase.data();
        d.surf_pbase = _pbase.data();
        d.nsurf = (uint32_t)_vbase.size() - 1;
        d.obj_surface = surf.data();
        d.obj_local2world = l2w.data();
        d.obj_world2local = w2l.data();
        d.obj_material = mat.data();
        d.nobj = (uint32_t)objects.size();
        d.lights = lt.data();
        d.nlight = (uint32_t)lights.size();
        for (int c = 0; c < 3; c++)
        {
            d.background[c] = backgroundColor[c];
            d.ambient[c] = ambientLight[c];
        }
        d.hash_seed = hashSeed;
        release();
        check(cgs_scene_create(device.handle(), &d, &_scene), "Scene::buildBVH");
    }

    cgs_scene* handle() const noexcept { return _scene; }

    Scene() = default;
    ~Scene() { release(); }
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;

private:
    void release() noexcept
    {
        if (_scene) (void)cgs_scene_destroy(_scene);
        _scene = nullptr;
    }
    std::vector<float> _vertices;
    std::vector<uint32_t> _indices;
    std::vector<uint32_t> _vbase {0}, _pbase {0};
    cgs_scene* _scene = nullptr;
};

// Packed-sRGB frame in device memory, row stride = width.
// rt::Frame (Frame.h:29-94) over caller-owned device memory: pixel (x, y) at
// data[stride*y + x]; stride 0 means the width.  alignTo gives the stride of
// a Frame built with an alignment (a power of two), as the reference pads.
struct Frame
{
    uint32_t* data;
    uint32_t width, height;
    uint32_t stride = 0;

    static constexpr uint32_t alignTo(uint32_t n, uint32_t alignment)
    {
        if (!alignment) alignment = 1;
        const uint32_t rem = n & (alignment - 1);
        return rem ? n + (alignment - rem) : n;
    }
};

// ---------------------------------------------------------------------------
// rt::RayTracer (rt/RayTracer.h:13-72).  The reference chooses the Monte Carlo
// render at compile time (USE_MONTECARLO_SAMPLING, rt/StaticConfig.h:5, off);
// here Options::monteCarlo chooses it per tracer (cgs_render_sampled, with
// Options::samples rays per pixel and draws from Options::seed).
// ---------------------------------------------------------------------------
class RayTracer
{
public:
    static constexpr float c_MinRadiance = 0x1p-8f;

    struct Options
    {
        int samples = 8;   // camera rays per pixel of the Monte Carlo render
        float eps = 1e-4f;
        bool flipYAxis = false;
        int recursionDepth = 6;
        int device = 0;
        bool monteCarlo = false;
        uint32_t seed = 42;
    };

    RayTracer() : RayTracer(Options {}) {}
    explicit RayTracer(Options&& op) : options(std::move(op)), _dev(options.device) {}

    // Asynchronous on `stream`.  Busy: warning and no-op (RayTracer.cu:109-115).
    void render(Frame* frame, const Camera* camera, const Scene* scene, stream_t stream = nullptr,
                cgs_ray_counts* deviceCounts = nullptr)
    {
        cgs_camera c {};
        for (int i = 0; i < 16; i++) c.to_world[i] = camera->toWorld[i];
        for (int i = 0; i < 3; i++) c.position[i] = camera->position[i];
        c.fovy = camera->fieldOfView;
        c.flip_y = options.flipYAxis ? 1 : 0;
        cgs_render_options o {options.eps, options.recursionDepth, frame->stride};
        cgs_status s = options.monteCarlo
                           ? cgs_render_sampled(_dev.handle(), scene->handle(), &c, &o, (uint32_t)options.samples,
                                                options.seed, frame->width, frame->height, frame->data, nullptr,
                                                deviceCounts, stream)
                           : cgs_render(_dev.handle(), scene->handle(), &c, &o, frame->width, frame->height,
                                        frame->data, nullptr, deviceCounts, stream);
        if (s == CGS_ERR_BUSY)
        {
            std::fprintf(stderr, "CUDA ray tracer not ready: wait the previous work to finish or create a new instance\n");
            return;
        }
        check(s, "RayTracer::render");
    }

    bool running() const { return cgs_render_running(_dev.handle()) != 0; }

    const Device& device() const noexcept { return _dev; }

    Options options;

private:
    Device _dev;
};

} // namespace rt

// ---------------------------------------------------------------------------
// Per-call forms of the reference's header functions
// ---------------------------------------------------------------------------

// Ray3f of the reference's Ds library: origin and direction (normalized where
// the reference's constructor normalizes it, by the library) and the range.
struct Ray3f
{
    float origin[3];
    float direction[3];
    float tMax = INFINITY;
};

// Intersection as the per-patch functions and PatchBVH fill it: distance
// (the current best on entry), the patch coordinates (u, v), the patch index
// and the object.
struct Intersection
{
    float distance = INFINITY;
    float p[2] = {0.0f, 0.0f};
    uint32_t triangleIndex = CGS_EMPTY;
    const void* object = nullptr;
};

namespace detail
{

// A small device staging area per GPU for the per-call functions (one call
// at a time per GPU; guarded).
struct Staging
{
    std::mutex mu;
    void* p = nullptr;
    static constexpr size_t bytes = 1024;
};
inline Staging& staging(const rt::Device& dev)
{
    static std::mutex mu;
    static std::map<const rt::Device*, Staging*> all;
    std::lock_guard lk(mu);
    Staging*& s = all[&dev];
    if (!s)
    {
        s = new Staging;
        s->p = dev.allocate(Staging::bytes);
    }
    return *s;
}

// One ray against one patch gathered from buffer[patch[k]] (vec4 points)
// by a batched entry point of size 1; hit is updated when found.
inline bool perPatch(decltype(&cgs_bezier_clip) fn, Intersection& hit, const Ray3f& ray, const float* buffer4,
                     const uint32_t patch[16], float param, const rt::Device& dev)
{
    cgs_ray r {ray.origin[0], ray.origin[1], ray.origin[2], hit.distance,
               ray.direction[0], ray.direction[1], ray.direction[2], 0.0f};
    float pts[64];
    for (int k = 0; k < 16; k++)
        for (int c = 0; c < 4; c++) pts[4 * k + c] = buffer4[4 * (size_t)patch[k] + c];
    Staging& st = staging(dev);
    std::lock_guard lk(st.mu);
    char* base = (char*)st.p;
    dev.upload(base, &r, sizeof r);
    dev.upload(base + 64, pts, sizeof pts);
    rt::check(fn(dev.handle(), (const cgs_ray*)base, (const float*)(base + 64), 1, param, (cgs_hit*)(base + 512), nullptr),
              "per-patch intersection");
    cgs_hit h;
    dev.download(&h, base + 512, sizeof h);
    dev.synchronize();
    if (h.patch == CGS_PATCH_OVERFLOW) throw rt::Error(CGS_ERR_STACK_OVERFLOW, "doBezierClipping");
    if (h.patch != 1u) return false;
    hit.distance = h.t;
    hit.p[0] = h.u;
    hit.p[1] = h.v;
    return true;
}

} // namespace detail

namespace spline
{

// spline::doBezierClipping (Spline.h:488-493) for one ray and one patch:
// buffer holds vec4 points (4 floats each), patch 16 indices into it.  The
// hit's distance is the current best; on a hit nearer than it, distance and
// p (u, v) are updated and true is returned.
inline bool doBezierClipping(Intersection& hit, const Ray3f& ray, const float* buffer4, const uint32_t patch[16],
                             float tol = 0x1p-12f, const rt::Device& dev = rt::Device::shared())
{
    return detail::perPatch(&cgs_bezier_clip, hit, ray, buffer4, patch, tol, dev);
}

// spline::doSubdivision (Spline.h:495-500) for one ray and one patch.
inline bool doSubdivision(Intersection& hit, const Ray3f& ray, const float* buffer4, const uint32_t patch[16],
                          float tol = 0x1p-4f, const rt::Device& dev = rt::Device::shared())
{
    return detail::perPatch(&cgs_subdivision, hit, ray, buffer4, patch, tol, dev);
}

// The header evaluations (Spline.h:197-254, Spline.cu:10-27,
// SplineMat.h:125-187), batched over device arrays: patches n × 16 float4,
// uv n × 2, out n × 4 (kind: CGS_EVAL_*).
inline void evaluate(uint32_t kind, const float* patches, const float* uv, uint32_t n, float* out,
                     const rt::Device& dev = rt::Device::shared(), rt::stream_t stream = nullptr)
{
    rt::check(cgs_patch_eval(dev.handle(), kind, patches, uv, n, out, stream), "spline evaluation");
}

// spline::subpatch (Spline.h:384-396), batched: ranges n × (umin, umax, vmin, vmax).
inline void subpatch(const float* patches, const float* ranges, uint32_t n, float* out,
                     const rt::Device& dev = rt::Device::shared(), rt::stream_t stream = nullptr)
{
    rt::check(cgs_patch_subpatch(dev.handle(), patches, ranges, n, out, stream), "spline::subpatch");
}

// One patch gathered from buffer[patch[k]] at (u, v) by `kind`, on the GPU:
// spline::interpolate (CGS_EVAL_INTERPOLATE), derivativeU/V, normal, and
// their mat:: forms.  Returns x, y, z, w (w = 0 for normals).
inline std::array<float, 4> evaluate(uint32_t kind, const float* buffer4, const uint32_t patch[16], float u, float v,
                                     const rt::Device& dev = rt::Device::shared())
{
    float pts[64];
    for (int k = 0; k < 16; k++)
        for (int c = 0; c < 4; c++) pts[4 * k + c] = buffer4[4 * (size_t)patch[k] + c];
    const float uv[2] = {u, v};
    detail::Staging& st = detail::staging(dev);
    std::lock_guard lk(st.mu);
    char* base = (char*)st.p;
    dev.upload(base, pts, sizeof pts);
    dev.upload(base + 256, uv, sizeof uv);
    evaluate(kind, (const float*)base, (const float*)(base + 256), 1, (float*)(base + 512), dev);
    std::array<float, 4> out;
    dev.download(out.data(), base + 512, sizeof(float) * 4);
    dev.synchronize();
    return out;
}

// spline::doBezierClipping (Spline.h:488-493), batched over device arrays:
// ray r against patch r (16 float4 control points each); hits[r].patch is 1
// on a hit, CGS_EMPTY on a miss.
inline void doBezierClipping(const rt::Device& dev, const cgs_ray* rays, const float* patches, uint32_t n,
                             cgs_hit* hits, float tol = 0x1p-12f, rt::stream_t stream = nullptr)
{
    rt::check(cgs_bezier_clip(dev.handle(), rays, patches, n, tol, hits, stream), "doBezierClipping");
}

// spline::doSubdivision (Spline.h:495-500), batched the same way.
inline void doSubdivision(const rt::Device& dev, const cgs_ray* rays, const float* patches, uint32_t n,
                          cgs_hit* hits, float threshold = 0x1p-4f, rt::stream_t stream = nullptr)
{
    rt::check(cgs_subdivision(dev.handle(), rays, patches, n, threshold, hits, stream), "doSubdivision");
}

} // namespace spline

// ---------------------------------------------------------------------------
// cg::PatchBVH (PatchBVH.h:13-28, PatchBVH.cpp:7-121): a BVH over one
// surface's Bézier patches, resident on a Device, with the queries of
// BVHBase that PatchBVH's leaves serve: the closest hit (intersect(ray, hit))
// and whether anything is hit (intersect(ray)).
// ---------------------------------------------------------------------------
class PatchBVH
{
public:
    explicit PatchBVH(const Surface* s, const rt::Device& dev = rt::Device::shared()) : _surface(s), _dev(&dev)
    {
        const auto pts = s->points();
        const auto idx = s->indices();
        const auto el = s->elements();
        if (el.empty())
            rt::check(cgs_patchbvh_create(dev.handle(), pts.data(), (uint32_t)(pts.size() / 4), idx.data(),
                                          (uint32_t)(idx.size() / 16), &_bvh),
                      "PatchBVH");
        else
            rt::check(cgs_patchbvh_create_elements(dev.handle(), pts.data(), (uint32_t)(pts.size() / 4), idx.data(),
                                                   (uint32_t)(idx.size() / 16), el.data(), &_bvh),
                      "PatchBVH");
    }
    ~PatchBVH()
    {
        if (_bvh) (void)cgs_patchbvh_destroy(_bvh);
    }
    PatchBVH(const PatchBVH&) = delete;
    PatchBVH& operator=(const PatchBVH&) = delete;

    const Surface* surface() const { return _surface; }

    // The closest patch hit nearer than hit.distance: fills distance, p (u, v),
    // triangleIndex (the patch) and object (this), as intersectLeaf does.
    bool intersect(const Ray3f& ray, Intersection& hit) const
    {
        cgs_ray r {ray.origin[0], ray.origin[1], ray.origin[2], std::min(hit.distance, ray.tMax),
                   ray.direction[0], ray.direction[1], ray.direction[2], 0.0f};
        cgs_hit h = one(r, false);
        if (h.patch == CGS_EMPTY) return false;
        if (h.patch == CGS_PATCH_OVERFLOW) throw rt::Error(CGS_ERR_STACK_OVERFLOW, "PatchBVH::intersect");
        hit.distance = h.t;
        hit.p[0] = h.u;
        hit.p[1] = h.v;
        hit.triangleIndex = h.patch;
        hit.object = this;
        return true;
    }

    // Whether a patch whose box the ray reaches within ray.tMax is hit.
    bool intersect(const Ray3f& ray) const
    {
        cgs_ray r {ray.origin[0], ray.origin[1], ray.origin[2], ray.tMax,
                   ray.direction[0], ray.direction[1], ray.direction[2], 0.0f};
        const cgs_hit h = one(r, true);
        if (h.patch == 2u) throw rt::Error(CGS_ERR_STACK_OVERFLOW, "PatchBVH::intersect");
        return h.patch == 1u;
    }

    // Batched over device arrays (n rays, n hits or n occlusion words).
    void intersect(const cgs_ray* rays, uint32_t n, cgs_hit* hits, rt::stream_t stream = nullptr) const
    {
        rt::check(cgs_patchbvh_intersect(_dev->handle(), _bvh, rays, n, hits, stream), "PatchBVH::intersect");
    }
    void intersectAny(const cgs_ray* rays, uint32_t n, uint32_t* occluded, rt::stream_t stream = nullptr) const
    {
        rt::check(cgs_patchbvh_intersect_any(_dev->handle(), _bvh, rays, n, occluded, stream), "PatchBVH::intersect");
    }

private:
    cgs_hit one(const cgs_ray& r, bool any) const
    {
        detail::Staging& st = detail::staging(*_dev);
        std::lock_guard lk(st.mu);
        char* base = (char*)st.p;
        _dev->upload(base, &r, sizeof r);
        if (any)
            intersectAny((const cgs_ray*)base, 1, (uint32_t*)(base + 512));
        else
            intersect((const cgs_ray*)base, 1, (cgs_hit*)(base + 512));
        cgs_hit h {};
        _dev->download(&h, base + 512, any ? sizeof(uint32_t) : sizeof h);
        _dev->synchronize();
        if (any) h.patch = *(const uint32_t*)&h.t;
        return h;
    }

    const Surface* _surface;
    const rt::Device* _dev;
    cgs_patchbvh* _bvh = nullptr;
};

} // namespace cg
