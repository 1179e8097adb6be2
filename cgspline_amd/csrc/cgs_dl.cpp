// SPDX-License-Identifier: MIT
//
// The reference's scene description language (`.dl` files; the grammar of
// reader/DescriptionLanguage.{h,cpp} and reader/Parser.cpp, the scene
// vocabulary of reader/SceneReader.cpp) read into the scene record that
// cgs_scene_create takes, flattened the way MainWindow::convertScene fills
// rt::Scene (MainWindow.cpp:363-492).
//
// Language.  A file is a sequence of statements `[let NAME =] value`:
//   value   := expr | dict | list | literal
//   expr    := root suffix*          root   := NAME | + | - | * | / | . NAME
//   suffix  := . NAME | ( value* ) | dict       (property access, call)
//   literal := ( n n [n [n]] ) | integer | float | "string"
//   dict    := { (NAME | "string") : value ... }     list := [ value* ]
// with // and /* */ comments.  A sign directly before a digit belongs to the
// number.  Built-in names: + - * / (functions of their arguments),
// From_sRGB, Assets, Get, Material, Surface, Scene, Mesh, Texture.
//
// Scene content.  Objects nest (`objects: [...]`); each may carry a
// `transform` {position, rotation (Euler degrees), scale}, a `light` and a
// `surface` {model, material}.  Flattening visits objects depth first, an
// object before its children, which is the order the reference's scene
// graph lists actors and lights.
//
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cgs_internal.hpp"

namespace
{

struct DlError : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

[[noreturn]] void fail(const std::string& msg) { throw DlError(msg); }

// ---------------------------------------------------------------------------
// values
// ---------------------------------------------------------------------------

struct Value;
using List = std::vector<Value>;
using Dict = std::map<std::string, Value>;
using Fn = std::function<Value(const List&)>;

struct Value
{
    enum Kind { None, Int, Float, Vec, Str, DictK, ListK, FnK, MaterialK, SurfaceK, MeshK, SceneK };
    Kind kind = None;
    int i = 0;
    float f = 0;
    int n = 0;        // vector size 2..4
    float v[4] {};
    std::string s;
    std::shared_ptr<Dict> d;
    std::shared_ptr<List> l;
    std::shared_ptr<Fn> fn;
    int asset = -1;   // material / surface index

    static Value ofInt(int x) { Value r; r.kind = Int; r.i = x; return r; }
    static Value ofFloat(float x) { Value r; r.kind = Float; r.f = x; return r; }
    static Value ofVec(int n, const float* p)
    {
        Value r;
        r.kind = Vec;
        r.n = n;
        for (int k = 0; k < n; ++k) r.v[k] = p[k];
        return r;
    }
    static Value ofStr(std::string x) { Value r; r.kind = Str; r.s = std::move(x); return r; }
    static Value ofFn(Fn f) { Value r; r.kind = FnK; r.fn = std::make_shared<Fn>(std::move(f)); return r; }
    static Value ofAsset(Kind k, int idx) { Value r; r.kind = k; r.asset = idx; return r; }

    bool scalar() const { return kind == Int || kind == Float; }
    float asFloat() const { return kind == Int ? float(i) : f; }

    const char* typeName() const
    {
        switch (kind)
        {
        case Int: return "int";
        case Float: return "float";
        case Vec: return n == 2 ? "vec2" : n == 3 ? "vec3" : "vec4";
        case Str: return "string";
        case FnK: return "function";
        case None: return "nothing";
        default: return "object";
        }
    }
};

[[noreturn]] void mismatch(const char* op, const Value& a, const Value& b)
{
    fail(std::string("operator `") + op + "` cannot be called with argument types `" + a.typeName() + "` and `" +
         b.typeName() + "`");
}

Value vecScale(const Value& a, float s)
{
    float p[4];
    for (int k = 0; k < a.n; ++k) p[k] = a.v[k] * s;
    return Value::ofVec(a.n, p);
}

// The four operators of DescriptionLanguage.cpp:27-235, over int, float and
// vectors of one size.
Value opAdd(const List& args)
{
    if (args.size() < 2) fail("expected at least 2 arguments");
    Value a = args[0];
    for (size_t k = 1; k < args.size(); ++k)
    {
        const Value& b = args[k];
        if (a.kind == Value::Int && b.kind == Value::Int) a = Value::ofInt(a.i + b.i);
        else if (a.scalar() && b.scalar()) a = Value::ofFloat(a.asFloat() + b.asFloat());
        else if (a.kind == Value::Vec && b.kind == Value::Vec && a.n == b.n)
        {
            float p[4];
            for (int c = 0; c < a.n; ++c) p[c] = a.v[c] + b.v[c];
            a = Value::ofVec(a.n, p);
        }
        else mismatch("+", a, b);
    }
    return a;
}

Value opSub(const List& args)
{
    if (args.size() == 1)
    {
        const Value& a = args[0];
        if (a.kind == Value::Int) return Value::ofInt(-a.i);
        if (a.kind == Value::Float) return Value::ofFloat(-a.f);
        if (a.kind == Value::Vec) return vecScale(a, -1.0f);
        fail(std::string("operator `-` not supported with type `") + a.typeName() + "`");
    }
    if (args.size() != 2) fail("expected one or two arguments");
    const Value &a = args[0], &b = args[1];
    if (a.kind == Value::Int && b.kind == Value::Int) return Value::ofInt(a.i - b.i);
    if (a.scalar() && b.scalar()) return Value::ofFloat(a.asFloat() - b.asFloat());
    if (a.kind == Value::Vec && b.kind == Value::Vec && a.n == b.n)
    {
        float p[4];
        for (int c = 0; c < a.n; ++c) p[c] = a.v[c] - b.v[c];
        return Value::ofVec(a.n, p);
    }
    mismatch("-", a, b);
}

Value opMul(const List& args)
{
    if (args.size() < 2) fail("expected at least 2 arguments");
    Value a = args[0];
    for (size_t k = 1; k < args.size(); ++k)
    {
        const Value& b = args[k];
        if (a.kind == Value::Int && b.kind == Value::Int) a = Value::ofInt(a.i * b.i);
        else if (a.scalar() && b.scalar()) a = Value::ofFloat(a.asFloat() * b.asFloat());
        else if (a.scalar() && b.kind == Value::Vec)
        {
            float p[4];
            for (int c = 0; c < b.n; ++c) p[c] = a.asFloat() * b.v[c];
            a = Value::ofVec(b.n, p);
        }
        else if (a.kind == Value::Vec && b.scalar()) a = vecScale(a, b.asFloat());
        else mismatch("*", a, b);
    }
    return a;
}

Value opDiv(const List& args)
{
    if (args.size() != 2) fail("expected only two arguments");
    const Value &a = args[0], &b = args[1];
    if (a.kind == Value::Int && b.kind == Value::Int)
    {
        if (b.i == 0) fail("integer division by zero");
        return Value::ofInt(a.i / b.i);
    }
    if (a.scalar() && b.scalar()) return Value::ofFloat(a.asFloat() / b.asFloat());
    // a scalar over a vector scales the vector by the scalar's reciprocal
    // (DescriptionLanguage.cpp:203-224)
    if (a.scalar() && b.kind == Value::Vec) return vecScale(b, 1.0f / a.asFloat());
    if (a.kind == Value::Vec && b.scalar()) return vecScale(a, 1.0f / b.asFloat());
    mismatch("/", a, b);
}

// ---------------------------------------------------------------------------
// tokens
// ---------------------------------------------------------------------------

enum : int { T_END = -1, T_ID = 256, T_LET, T_INT, T_FLOAT, T_STR };

struct Token
{
    int kind;
    std::string text;
    int ival = 0;
    float fval = 0;
    int line, col;
};

std::vector<Token> tokenize(const std::string& src, const std::string& name)
{
    std::vector<Token> out;
    size_t p = 0;
    int line = 1, col = 1;
    auto at = [&](size_t q) -> int { return q < src.size() ? (unsigned char)src[q] : -1; };
    auto adv = [&]() {
        if (src[p] == '\n') ++line, col = 1;
        else ++col;
        ++p;
    };
    auto where = [&](int l, int c) { return name + ":" + std::to_string(l) + ":" + std::to_string(c) + ": "; };
    for (;;)
    {
        while (p < src.size() && std::isspace((unsigned char)src[p])) adv();
        if (p >= src.size()) break;
        const int l0 = line, c0 = col;
        const int c = at(p);
        if (c == '/' && at(p + 1) == '/')
        {
            while (p < src.size() && src[p] != '\n') adv();
            continue;
        }
        if (c == '/' && at(p + 1) == '*')
        {
            adv(), adv();
            while (p < src.size() && !(src[p] == '*' && at(p + 1) == '/')) adv();
            if (p >= src.size()) fail(where(l0, c0) + "unexpected end-of-file");
            adv(), adv();
            continue;
        }
        Token t {0, "", 0, 0, l0, c0};
        if (std::isalpha(c) || c == '_')
        {
            while (p < src.size() && (std::isalnum((unsigned char)src[p]) || src[p] == '_')) t.text += src[p], adv();
            t.kind = t.text == "let" ? T_LET : T_ID;
        }
        else if (std::isdigit(c) || ((c == '+' || c == '-') && std::isdigit(at(p + 1))))
        {
            bool isFloat = false;
            t.text += src[p], adv();
            while (std::isdigit(at(p))) t.text += src[p], adv();
            if (at(p) == '.')
            {
                isFloat = true;
                t.text += src[p], adv();
                while (std::isdigit(at(p))) t.text += src[p], adv();
            }
            if (at(p) == 'e' || at(p) == 'E')
            {
                isFloat = true;
                t.text += src[p], adv();
                if (at(p) == '+' || at(p) == '-') t.text += src[p], adv();
                if (!std::isdigit(at(p))) fail(where(line, col) + "malformed number exponent");
                while (std::isdigit(at(p))) t.text += src[p], adv();
            }
            errno = 0;
            if (isFloat)
            {
                t.kind = T_FLOAT;
                t.fval = std::strtof(t.text.c_str(), nullptr);
                if (errno == ERANGE && std::isinf(t.fval)) fail(where(l0, c0) + "float out of range: " + t.text);
            }
            else
            {
                t.kind = T_INT;
                const long long x = std::strtoll(t.text.c_str(), nullptr, 10);
                if (errno == ERANGE || x > 2147483647LL || x < -2147483648LL)
                    fail(where(l0, c0) + "integer out of range: " + t.text);
                t.ival = int(x);
            }
        }
        else if (c == '"')
        {
            adv();
            while (at(p) != '"')
            {
                if (at(p) < 0 || at(p) == '\n') fail(where(l0, c0) + "missing string literal ending");
                t.text += src[p], adv();
            }
            adv();
            t.kind = T_STR;
        }
        else
        {
            t.kind = c;
            t.text = std::string(1, char(c));
            adv();
        }
        out.push_back(std::move(t));
    }
    out.push_back(Token {T_END, "end of file", 0, 0, line, col});
    return out;
}

// ---------------------------------------------------------------------------
// scene content collected while reading
// ---------------------------------------------------------------------------

struct Material
{
    float diffuse[3], specular[3], metalness, roughness;
};

// Defaults where a scene file leaves a value out.  The reference takes them
// from its graphics library, which is not part of the reference checkout:
// a material made by `Material {}` starts black (SceneReader.cpp:57 passes
// Color::black), a surface without a material gets a neutral grey.
constexpr Material kBlackMaterial {{0, 0, 0}, {0, 0, 0}, 0.0f, 1.0f};
constexpr Material kDefaultMaterial {{0.8f, 0.8f, 0.8f}, {0.04f, 0.04f, 0.04f}, 0.0f, 0.5f};

struct SurfaceAsset
{
    std::string name;
    std::vector<float> vertices;     // 4 per point
    std::vector<uint32_t> indices;   // 16 per patch
};

struct Mat4
{
    double m[4][4];   // m[row][col]
    static Mat4 identity()
    {
        Mat4 r {};
        for (int k = 0; k < 4; ++k) r.m[k][k] = 1;
        return r;
    }
    Mat4 operator*(const Mat4& b) const
    {
        Mat4 r {};
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) r.m[i][j] += m[i][k] * b.m[k][j];
        return r;
    }
    void store(float* colMajor) const
    {
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) colMajor[4 * c + r] = float(m[r][c]);
    }
    // inverse of an affine transform [A t; 0 1]
    Mat4 inverseAffine() const
    {
        const double a = m[0][0], b = m[0][1], c = m[0][2], d = m[1][0], e = m[1][1], f = m[1][2], g = m[2][0],
                     h = m[2][1], i = m[2][2];
        const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
        const double det = a * A + b * B + c * C;
        if (det == 0 || !std::isfinite(det)) fail("transform is not invertible (zero scale)");
        Mat4 r = identity();
        const double inv[3][3] = {{A, -(b * i - c * h), b * f - c * e},
                                  {B, a * i - c * g, -(a * f - c * d)},
                                  {C, -(a * h - b * g), a * e - b * d}};
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) r.m[x][y] = inv[x][y] / det;
        for (int x = 0; x < 3; ++x)
            r.m[x][3] = -(r.m[x][0] * m[0][3] + r.m[x][1] * m[1][3] + r.m[x][2] * m[2][3]);
        return r;
    }
};

// R = Ry(y)·Rx(x)·Rz(z) for Euler angles in degrees: yaw, then pitch, then
// roll.  The reference's graphics library defines the order; it is not in
// the checkout, so this order is this reader's choice.
Mat4 eulerRotation(const float* deg)
{
    const double k = 3.14159265358979323846 / 180.0;
    const double x = deg[0] * k, y = deg[1] * k, z = deg[2] * k;
    Mat4 Rx = Mat4::identity(), Ry = Mat4::identity(), Rz = Mat4::identity();
    Rx.m[1][1] = std::cos(x), Rx.m[1][2] = -std::sin(x), Rx.m[2][1] = std::sin(x), Rx.m[2][2] = std::cos(x);
    Ry.m[0][0] = std::cos(y), Ry.m[0][2] = std::sin(y), Ry.m[2][0] = -std::sin(y), Ry.m[2][2] = std::cos(y);
    Rz.m[0][0] = std::cos(z), Rz.m[0][1] = -std::sin(z), Rz.m[1][0] = std::sin(z), Rz.m[1][1] = std::cos(z);
    return Ry * Rx * Rz;
}

struct FlatObject
{
    std::string name;
    int surface;     // surface asset index; -1: a mesh (no geometry on this path)
    Material material;
    Mat4 world;
};

struct FlatLight
{
    std::string name;
    float row[12];
};

struct SceneState
{
    bool present = false;
    float background[3] {0.1f, 0.1f, 0.1f};
    float ambient[3] {0.1f, 0.1f, 0.1f};
    float viewPosition[3] {0, 0, 0};
    float viewRotation[3] {0, 0, 0};
    std::vector<FlatObject> objects;
    std::vector<FlatLight> lights;
};

// ---------------------------------------------------------------------------
// reader
// ---------------------------------------------------------------------------

class Reader
{
public:
    Reader(std::string name, std::string workingDir) : _name(std::move(name)), _dir(std::move(workingDir))
    {
        _assets = std::make_shared<Dict>();
        Value assets;
        assets.kind = Value::DictK;
        assets.d = _assets;
        _scope["Assets"] = assets;
        _scope["+"] = Value::ofFn(opAdd);
        _scope["-"] = Value::ofFn(opSub);
        _scope["*"] = Value::ofFn(opMul);
        _scope["/"] = Value::ofFn(opDiv);
        // DescriptionLanguage.cpp:247 tests `n != 3 || n != 4`, which
        // always holds: From_sRGB fails for every argument list.
        _scope["From_sRGB"] = Value::ofFn([](const List&) -> Value {
            fail("From_sRGB function: expected 3 or 4 integer arguments");
        });
        _scope["Get"] = Value::ofFn([this](const List& a) { return get(a); });
        _scope["Material"] = Value::ofFn([this](const List& a) { return createMaterial(a); });
        _scope["Surface"] = Value::ofFn([this](const List& a) { return createSurface(a); });
        _scope["Scene"] = Value::ofFn([this](const List& a) { return readScene(a); });
        _scope["Mesh"] = Value::ofFn([this](const List& a) { return createMesh(a); });
        // the reference's built-in box mesh (SceneReader.cpp:40)
        insertAsset("Box", Value::ofAsset(Value::MeshK, 0));
        _scope["Texture"] = Value::ofFn([](const List&) -> Value { fail("Texture: textures are not supported"); });
    }

    void parse(const std::string& text)
    {
        _tok = tokenize(text, _name);
        _p = 0;
        while (cur().kind != T_END) statement();
        if (!_scene.present) fail(_name + ": the file defines no Scene");
    }

    std::vector<Material> materials;
    std::vector<SurfaceAsset> surfaces;
    SceneState _scene;

private:
    std::string _name, _dir;
    std::vector<Token> _tok;
    size_t _p = 0;
    Dict _scope;
    std::shared_ptr<Dict> _assets;
    int _materialId = 0, _surfaceId = 0, _meshId = 0, _objectId = 0;

    const Token& cur() const { return _tok[_p]; }
    std::string where(const Token& t) const
    {
        return _name + ":" + std::to_string(t.line) + ":" + std::to_string(t.col) + ": ";
    }
    [[noreturn]] void error(const std::string& msg) const { fail(where(cur()) + msg); }
    void next() { if (cur().kind != T_END) ++_p; }
    void match(int kind)
    {
        if (cur().kind != kind)
            error(std::string("expected '") + (kind < 256 ? std::string(1, char(kind)) : std::string("token")) +
                  "', found '" + cur().text + "'");
        next();
    }
    std::string identifier()
    {
        if (cur().kind != T_ID) error("identifier expected, found '" + cur().text + "'");
        std::string s = cur().text;
        next();
        return s;
    }

    // Statement <- (LET IDENTIFIER '=')? Value
    void statement()
    {
        if (cur().kind == T_LET)
        {
            next();
            const Token at = cur();
            std::string id = identifier();
            if (_scope.count(id)) fail(where(at) + "identifier \"" + id + "\" already defined");
            match('=');
            Value v = value();
            _scope[id] = std::move(v);
            return;
        }
        (void)value();
    }

    static bool isOperator(int k) { return k == '+' || k == '-' || k == '*' || k == '/'; }

    // Value <- Expr / Dict / List / Literal
    // Lists, dictionaries and call arguments nest through value(); a file
    // nested deeper than this is refused rather than exhausting the stack.
    static constexpr int kMaxNesting = 256;
    int _nesting = 0;

    Value value()
    {
        struct Leave
        {
            int& n;
            ~Leave() { --n; }
        } leave {++_nesting};
        if (_nesting > kMaxNesting) error("values nested more than " + std::to_string(kMaxNesting) + " deep");
        const int k = cur().kind;
        if (k == T_ID || isOperator(k) || k == '.') return expr();
        if (k == '{') return dict();
        if (k == '[') return list();
        return literal();
    }

    Value literal()
    {
        const Token& t = cur();
        if (t.kind == '(') return vector();
        if (t.kind == T_INT) { Value v = Value::ofInt(t.ival); next(); return v; }
        if (t.kind == T_FLOAT) { Value v = Value::ofFloat(t.fval); next(); return v; }
        if (t.kind == T_STR) { Value v = Value::ofStr(t.text); next(); return v; }
        error("expected a constant literal value, found '" + t.text + "'");
    }

    float number()
    {
        const Token& t = cur();
        if (t.kind != T_INT && t.kind != T_FLOAT) error("a number was expected, found '" + t.text + "'");
        const float f = t.kind == T_INT ? float(t.ival) : t.fval;
        next();
        return f;
    }

    // Vector <- '(' n n [n [n]] ')'
    Value vector()
    {
        match('(');
        float p[4];
        int n = 0;
        p[n++] = number();
        p[n++] = number();
        while (n < 4 && (cur().kind == T_INT || cur().kind == T_FLOAT)) p[n++] = number();
        match(')');
        return Value::ofVec(n, p);
    }

    Value list()
    {
        Value r;
        r.kind = Value::ListK;
        r.l = std::make_shared<List>();
        match('[');
        while (cur().kind != ']')
        {
            if (cur().kind == T_END) error("unexpected end-of-file in a list");
            r.l->push_back(value());
        }
        match(']');
        return r;
    }

    // Dict <- '{' ((IDENTIFIER / STRING) ':' Value)* '}'
    Value dict()
    {
        Value r;
        r.kind = Value::DictK;
        r.d = std::make_shared<Dict>();
        match('{');
        while (cur().kind == T_ID || cur().kind == T_STR)
        {
            const Token at = cur();
            std::string key = cur().text;
            next();
            match(':');
            Value v = value();
            if (r.d->count(key)) fail(where(at) + "property \"" + key + "\" already defined");
            (*r.d)[key] = std::move(v);
        }
        if (cur().kind != '}') error("property definition expected, found '" + cur().text + "'");
        next();
        return r;
    }

    Value access(const Dict& d, const std::string& key, const Token& at) const
    {
        auto it = d.find(key);
        if (it == d.end()) fail(where(at) + "property \"" + key + "\" undefined");
        return it->second;
    }

    // Expr <- Root Suffix*
    Value expr()
    {
        Value v;
        const Token at = cur();
        if (cur().kind == T_ID) v = access(_scope, identifier(), at);
        else if (isOperator(cur().kind))
        {
            const std::string op = cur().text;
            next();
            v = access(_scope, op, at);
        }
        else v = property(_scope);
        for (;;)
        {
            if (cur().kind == '.')
            {
                if (v.kind != Value::DictK)
                    error("expression result is not a dictionary that holds any property");
                v = property(*v.d);
            }
            else if (cur().kind == '(' || cur().kind == '{')
            {
                if (v.kind != Value::FnK) error("expression result is not a function");
                v = call(*v.fn);
            }
            else return v;
        }
    }

    // PropertyAccess <- '.' (IDENTIFIER / STRING)
    Value property(const Dict& d)
    {
        match('.');
        const Token at = cur();
        if (at.kind != T_ID && at.kind != T_STR) error("a property name was expected");
        next();
        return access(d, at.text, at);
    }

    // FunctionCall <- '(' Value* ')' / Dict
    Value call(const Fn& f)
    {
        const Token at = cur();
        List args;
        if (cur().kind == '{') args.push_back(dict());
        else
        {
            match('(');
            while (cur().kind != ')')
            {
                if (cur().kind == T_END) error("unexpected end-of-file in an argument list");
                args.push_back(value());
            }
            next();
        }
        try
        {
            return f(args);
        }
        catch (const DlError& e)
        {
            fail(where(at) + "function evaluation failed: " + e.what());
        }
    }

    // --- property helpers ------------------------------------------------

    static const Value* prop(const Dict& d, const char* key)
    {
        auto it = d.find(key);
        return it == d.end() ? nullptr : &it->second;
    }
    static const Dict& asDict(const Value& v, const char* what)
    {
        if (v.kind != Value::DictK) fail(std::string(what) + ": expected a dictionary, got " + v.typeName());
        return *v.d;
    }
    static const std::string& asString(const Value& v, const char* what)
    {
        if (v.kind != Value::Str) fail(std::string(what) + ": expected a string, got " + v.typeName());
        return v.s;
    }
    static float asNumber(const Value& v, const char* what)
    {
        if (!v.scalar()) fail(std::string(what) + ": expected a number, got " + v.typeName());
        return v.asFloat();
    }
    static void asVec3(const Value& v, const char* what, float* out)
    {
        if (v.kind != Value::Vec || v.n != 3) fail(std::string(what) + ": expected a vec3, got " + v.typeName());
        for (int k = 0; k < 3; ++k) out[k] = v.v[k];
    }
    // a colour is a vec3 or a vec4 (its alpha dropped, as rt::Scene keeps rgb)
    static void asColor(const Value& v, const char* what, float* out)
    {
        if (v.kind != Value::Vec || v.n < 3)
            fail(std::string(what) + ": expected a colour (vec3 or vec4), got " + v.typeName());
        for (int k = 0; k < 3; ++k) out[k] = v.v[k];
    }

    void insertAsset(const std::string& name, const Value& v)
    {
        if (_assets->count(name)) fail("asset \"" + name + "\" already exists");
        (*_assets)[name] = v;
    }

    // --- built-in functions (SceneReader.cpp) ------------------------------

    Value get(const List& a)
    {
        if (a.empty()) fail("Get: expected an asset name");
        const std::string& name = asString(a[0], "Get");
        auto it = _assets->find(name);
        if (it == _assets->end()) fail("property \"" + name + "\" does not exist");
        return it->second;
    }

    // SceneReader.cpp:54-118
    Value createMaterial(const List& a)
    {
        if (a.empty()) fail("Material: expected a dictionary");
        const Dict& d = asDict(a[0], "Material");
        Material m = kBlackMaterial;
        std::string name = "Material_" + std::to_string(_materialId + 1);
        if (auto p = prop(d, "name")) name = asString(*p, "Material: name");
        else ++_materialId;
        if (auto p = prop(d, "ambient")) { float t[3]; asColor(*p, "Material: ambient", t); }
        if (auto p = prop(d, "diffuse")) asColor(*p, "Material: diffuse", m.diffuse);
        if (auto p = prop(d, "specular")) asColor(*p, "Material: specular", m.specular);
        if (auto p = prop(d, "transparency")) { float t[3]; asColor(*p, "Material: transparency", t); }
        if (auto p = prop(d, "refractive_index"))
            if (asNumber(*p, "Material: refractive_index") <= 0)
                fail("Material: 'refractive_index' (index of refraction) should be greater than zero");
        if (auto p = prop(d, "metalness"))
        {
            const float v = asNumber(*p, "Material: metalness");
            if (v < 0 || v > 1) fail("Material: 'metalness' should be between 0 and 1");
            m.metalness = v;
        }
        if (auto p = prop(d, "roughness"))
        {
            const float v = asNumber(*p, "Material: roughness");
            if (v < 0 || v > 1) fail("Material: 'roughness' should be between 0 and 1");
            m.roughness = v;
        }
        if (prop(d, "texture") || prop(d, "metal_rough_texture")) fail("Material: textures are not supported");
        materials.push_back(m);
        const Value v = Value::ofAsset(Value::MaterialK, int(materials.size()) - 1);
        insertAsset(name, v);
        return v;
    }

    // SceneReader.cpp:215-242 and Surface::load (Surface.cpp:350-373): a
    // name ending in ".se" is a `.se` surface (cgs_load_se), anything else a
    // `.be` model.
    Value createSurface(const List& a)
    {
        if (a.empty()) fail("Surface: expected a dictionary");
        const Dict& d = asDict(a[0], "Surface");
        std::string name = "Surface_" + std::to_string(_surfaceId + 1);
        if (auto p = prop(d, "name")) name = asString(*p, "Surface: name");
        else ++_surfaceId;
        const Value* file = prop(d, "file");
        if (!file) fail("surface data or source not provided");
        const std::string& rel = asString(*file, "Surface: file");
        const std::string path = rel.size() && rel[0] == '/' ? rel : (_dir.empty() ? rel : _dir + "/" + rel);
        std::string ext = rel.size() >= 3 ? rel.substr(rel.size() - 3) : "";
        for (char& c : ext) c = char(std::tolower((unsigned char)c));
        FILE* fp = std::fopen(path.c_str(), "rb");
        if (!fp) fail("file '" + rel + "' does not exist");
        std::string text;
        char buf[65536];
        for (size_t got; (got = std::fread(buf, 1, sizeof buf, fp)) > 0;) text.append(buf, got);
        std::fclose(fp);
        if (ext == ".se")
        {
            // MainWindow::convertScene (MainWindow.cpp:363-492) copies the
            // surface's points and its whole grouped index list into a
            // BezierSurface, which reads indexCount / 16 patches of 16
            // consecutive indices (Primitive.h:110): the elements' own node
            // counts and basis matrices play no part in the render.
            cgs_patch_table* t = nullptr;
            if (cgs_load_se(text.data(), text.size(), &t) != CGS_OK) fail("'" + rel + "': " + cgs_last_error());
            SurfaceAsset s;
            s.name = name;
            s.vertices.assign(t->points, t->points + 4 * size_t(t->npoints));
            const size_t whole = size_t(t->nindices) / 16 * 16;
            s.indices.assign(t->indices, t->indices + whole);
            cgs_patch_table_free(t);
            if (whole == 0) fail("'" + rel + "': fewer than 16 node indices, no patch");
            surfaces.push_back(std::move(s));
            const Value r = Value::ofAsset(Value::SurfaceK, int(surfaces.size()) - 1);
            insertAsset(name, r);
            return r;
        }
        float* v = nullptr;
        uint32_t* ix = nullptr;
        uint32_t nv = 0, np = 0;
        if (cgs_load_be(text.data(), text.size(), &v, &nv, &ix, &np) != CGS_OK)
            fail("'" + rel + "': " + cgs_last_error());
        SurfaceAsset s;
        s.name = name;
        s.vertices.assign(v, v + 4 * size_t(nv));
        s.indices.assign(ix, ix + 16 * size_t(np));
        cgs_free_host(v);
        cgs_free_host(ix);
        surfaces.push_back(std::move(s));
        const Value r = Value::ofAsset(Value::SurfaceK, int(surfaces.size()) - 1);
        insertAsset(name, r);
        return r;
    }

    // SceneReader.cpp:133-213: a triangle mesh from `vertices` (3 numbers per
    // vertex), optional `normals` (as many) and integer `triangles` (3 per
    // triangle).  Meshes have no geometry on this path (the reference gives
    // them empty bounds, Primitive.cu:47-56, so no ray reaches them); the
    // asset is checked and kept by name only.
    Value createMesh(const List& a)
    {
        if (a.empty()) fail("Mesh: expected a dictionary");
        const Dict& d = asDict(a[0], "Mesh");
        std::string name = "Mesh_" + std::to_string(_meshId + 1);
        if (auto p = prop(d, "name")) name = asString(*p, "Mesh: name");
        else ++_meshId;
        if (prop(d, "file")) fail("not yet implemented");
        size_t nv = 0;
        if (auto p = prop(d, "vertices"))
        {
            if (p->kind != Value::ListK) fail("Mesh: vertices must be a list");
            for (const Value& x : *p->l) (void)asNumber(x, "Mesh: vertices");
            nv = p->l->size() / 3;
        }
        if (auto p = prop(d, "normals"))
        {
            if (p->kind != Value::ListK) fail("Mesh: normals must be a list");
            if (p->l->size() / 3 != nv) fail("`normals` size must match `vertices` size");
            for (const Value& x : *p->l) (void)asNumber(x, "Mesh: normals");
        }
        if (auto p = prop(d, "triangles"))
        {
            if (p->kind != Value::ListK) fail("Mesh: triangles must be a list");
            for (const Value& x : *p->l)
                if (x.kind != Value::Int) fail("Mesh: triangle indices must be integers");
        }
        const Value v = Value::ofAsset(Value::MeshK, 0);
        insertAsset(name, v);
        return v;
    }

    // SceneReader.cpp:244-263 (addMesh): `name` is required and names the
    // mesh asset unless `model` gives it.
    FlatObject meshObject(const Dict& d, const Mat4& world, const std::string& objectName)
    {
        const Value* nm = prop(d, "name");
        if (!nm) fail("property \"name\" does not exist");
        const std::string& name = asString(*nm, "mesh: name");
        const Value* model = prop(d, "model");
        if (!model || model->kind != Value::MeshK)
        {
            auto it = _assets->find(name);
            if (it == _assets->end()) fail("property \"" + name + "\" does not exist");
            if (it->second.kind != Value::MeshK) fail("mesh: `" + name + "` is not a Mesh");
        }
        Material m = kDefaultMaterial;
        if (auto q = prop(d, "material"))
        {
            if (q->kind != Value::MaterialK) fail("mesh: `material` is not a Material");
            m = materials[q->asset];
        }
        return FlatObject {objectName, -1, m, world};
    }

    // SceneReader.cpp:420-456
    Value readScene(const List& a)
    {
        if (a.empty()) fail("Scene: expected a dictionary");
        const Dict& d = asDict(a[0], "Scene");
        _scene = SceneState {};
        _scene.present = true;
        if (auto p = prop(d, "environment"))
        {
            const Dict& e = asDict(*p, "Scene: environment");
            if (auto q = prop(e, "ambient")) asColor(*q, "environment: ambient", _scene.ambient);
            if (auto q = prop(e, "background")) asColor(*q, "environment: background", _scene.background);
            if (prop(e, "texture")) fail("environment: textures are not supported");
        }
        if (auto p = prop(d, "view"))
        {
            const Dict& v = asDict(*p, "Scene: view");
            if (auto q = prop(v, "position")) asVec3(*q, "view: position", _scene.viewPosition);
            if (auto q = prop(v, "rotation")) asVec3(*q, "view: rotation", _scene.viewRotation);
        }
        if (auto p = prop(d, "objects"))
        {
            if (p->kind != Value::ListK) fail("Scene: objects must be a list");
            for (const Value& o : *p->l) sceneObject(asDict(o, "Scene: object"), Mat4::identity());
        }
        Value r;
        r.kind = Value::SceneK;
        return r;
    }

    static Mat4 transform(const Dict& t)
    {
        float pos[3] {0, 0, 0}, rot[3] {0, 0, 0}, scale[3] {1, 1, 1};
        if (auto p = prop(t, "position")) asVec3(*p, "transform: position", pos);
        if (auto p = prop(t, "rotation")) asVec3(*p, "transform: rotation", rot);
        if (auto p = prop(t, "scale")) asVec3(*p, "transform: scale", scale);
        Mat4 T = Mat4::identity(), S = Mat4::identity();
        for (int k = 0; k < 3; ++k) T.m[k][3] = pos[k], S.m[k][k] = scale[k];
        return T * eulerRotation(rot) * S;
    }

    // SceneReader.cpp:328-374 (createLight) with MainWindow.cpp:375-404
    // (the rt::Light it becomes).  A light without a type is a point light.
    static FlatLight light(const Dict& d, const Mat4& world, const std::string& name)
    {
        enum { Directional, Point, Spot } type = Point;
        float color[3] {1, 1, 1}, range = 0, spotDeg = 45;
        if (auto p = prop(d, "type"))
        {
            const std::string& s = asString(*p, "Light: type");
            if (s == "point") type = Point;
            else if (s == "directional") type = Directional;
            else if (s == "spot") type = Spot;
            else fail("Light: type must be one of: {point, directional, spot}");
        }
        if (auto p = prop(d, "angle"))
        {
            spotDeg = asNumber(*p, "Light: angle");
            if (!(spotDeg > 0)) fail("Light: angle property must be positive");
        }
        if (auto p = prop(d, "color")) asColor(*p, "Light: color", color);
        if (auto p = prop(d, "range"))
        {
            range = asNumber(*p, "Light: range");
            if (!(range > 0)) fail("Light: range property must be positive");
        }
        if (auto p = prop(d, "falloff"))
            if (p->kind != Value::Int || p->i < 0 || p->i > 2)
                fail("Light: falloff property must be an integer in [0,2]");
        FlatLight L;
        L.name = name;
        float* r = L.row;
        r[0] = color[0], r[1] = color[1], r[2] = color[2], r[3] = 1.0f;
        for (int k = 0; k < 3; ++k) r[4 + k] = float(world.m[k][3]);
        r[7] = range;
        // direction: the object's −Z axis in world space
        double dir[3] = {-world.m[0][2], -world.m[1][2], -world.m[2][2]};
        const double len = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        for (int k = 0; k < 3; ++k) r[8 + k] = float(len > 0 ? dir[k] / len : (k == 2 ? -1.0 : 0.0));
        r[11] = type == Directional ? 0.0f : type == Point ? 3.1415f : float(spotDeg * (3.14159265358979323846 / 180.0));
        return L;
    }

    // SceneReader.cpp:387-418
    void sceneObject(const Dict& d, const Mat4& parent)
    {
        std::string name;
        if (auto p = prop(d, "name")) name = asString(*p, "object: name");
        else name = "Object " + std::to_string(++_objectId);
        Mat4 world = parent;
        if (auto p = prop(d, "transform")) world = parent * transform(asDict(*p, "object: transform"));
        if (auto p = prop(d, "light")) _scene.lights.push_back(light(asDict(*p, "object: light"), world, name));
        if (auto p = prop(d, "mesh")) _scene.objects.push_back(meshObject(asDict(*p, "object: mesh"), world, name));
        if (auto p = prop(d, "surface"))
        {
            const Dict& s = asDict(*p, "object: surface");
            const Value* model = prop(s, "model");
            if (!model) fail("property \"model\" does not exist");
            if (model->kind != Value::SurfaceK) fail("surface: `model` is not a Surface");
            Material m = kDefaultMaterial;
            if (auto q = prop(s, "material"))
            {
                if (q->kind != Value::MaterialK) fail("invalid material type for surface");
                m = materials[q->asset];
            }
            _scene.objects.push_back({name, model->asset, m, world});
        }
        if (auto p = prop(d, "objects"))
        {
            if (p->kind != Value::ListK) fail("object: objects must be a list");
            for (const Value& o : *p->l) sceneObject(asDict(o, "object"), world);
        }
    }
};

// The record handed to the caller and the storage its pointers point into
// (pub.storage points back at the owner).
struct DlScene
{
    cgs_dl_scene pub;
    std::vector<float> vertices, l2w, w2l, material, lights;
    std::vector<uint32_t> indices, vbase, pbase, objSurface;
    std::vector<std::string> objectNames, lightNames;
    std::vector<const char*> objectNamePtrs, lightNamePtrs;
};

DlScene* flatten(Reader& r)
{
    auto out = std::make_unique<DlScene>();
    const SceneState& sc = r._scene;
    // surfaces in order of first use; objects share a surface they name alike
    std::vector<int> slot(r.surfaces.size(), -1);
    out->vbase.push_back(0);
    out->pbase.push_back(0);
    for (const FlatObject& o : sc.objects)
    {
        if (o.surface >= 0 && slot[o.surface] < 0)
        {
            const SurfaceAsset& s = r.surfaces[o.surface];
            slot[o.surface] = int(out->vbase.size()) - 1;
            out->vertices.insert(out->vertices.end(), s.vertices.begin(), s.vertices.end());
            out->indices.insert(out->indices.end(), s.indices.begin(), s.indices.end());
            out->vbase.push_back(out->vbase.back() + uint32_t(s.vertices.size() / 4));
            out->pbase.push_back(out->pbase.back() + uint32_t(s.indices.size() / 16));
        }
        // a mesh object has no surface (CGS_EMPTY: empty bounds, never hit)
        out->objSurface.push_back(o.surface < 0 ? CGS_EMPTY : uint32_t(slot[o.surface]));
        float m[16];
        o.world.store(m);
        out->l2w.insert(out->l2w.end(), m, m + 16);
        o.world.inverseAffine().store(m);
        out->w2l.insert(out->w2l.end(), m, m + 16);
        const Material& q = o.material;
        const float mat[8] = {q.diffuse[0], q.diffuse[1], q.diffuse[2], q.specular[0], q.specular[1], q.specular[2],
                              q.metalness, q.roughness};
        out->material.insert(out->material.end(), mat, mat + 8);
        out->objectNames.push_back(o.name);
    }
    for (const FlatLight& L : sc.lights)
    {
        out->lights.insert(out->lights.end(), L.row, L.row + 12);
        out->lightNames.push_back(L.name);
    }
    for (auto& s : out->objectNames) out->objectNamePtrs.push_back(s.c_str());
    for (auto& s : out->lightNames) out->lightNamePtrs.push_back(s.c_str());

    cgs_dl_scene& p = out->pub;
    std::memset(&p, 0, sizeof p);
    cgs_scene_desc& d = p.scene;
    d.vertices = out->vertices.data();
    d.indices = out->indices.data();
    d.surf_vbase = out->vbase.data();
    d.surf_pbase = out->pbase.data();
    d.nsurf = uint32_t(out->vbase.size() - 1);
    d.obj_surface = out->objSurface.data();
    d.obj_local2world = out->l2w.data();
    d.obj_world2local = out->w2l.data();
    d.obj_material = out->material.data();
    d.nobj = uint32_t(sc.objects.size());
    d.lights = out->lights.data();
    d.nlight = uint32_t(sc.lights.size());
    for (int k = 0; k < 3; ++k) d.background[k] = sc.background[k], d.ambient[k] = sc.ambient[k];
    d.hash_seed = 42;
    Mat4 view = eulerRotation(sc.viewRotation);
    for (int k = 0; k < 3; ++k) view.m[k][3] = sc.viewPosition[k], p.view_position[k] = sc.viewPosition[k];
    view.store(p.view_to_world);
    p.object_names = out->objectNamePtrs.data();
    p.light_names = out->lightNamePtrs.data();
    p.storage = out.get();
    return out.release();
}

cgs_status parseInto(const std::string& text, const std::string& name, const std::string& dir, cgs_dl_scene** out)
{
    try
    {
        Reader r(name, dir);
        r.parse(text);
        bool anySurface = false;
        for (const FlatObject& o : r._scene.objects) anySurface |= o.surface >= 0;
        if (!anySurface) fail(name + ": the scene has no surface objects");
        *out = &flatten(r)->pub;
        return CGS_OK;
    }
    catch (const DlError& e)
    {
        return cgs::setError(CGS_ERR_INVALID, e.what());
    }
    catch (const std::bad_alloc&)
    {
        return cgs::setError(CGS_ERR_OUT_OF_MEMORY, "cgs_dl: out of host memory");
    }
}

} // namespace

cgs_status cgs_dl_parse(const char* text, size_t length, const char* name, const char* working_dir,
                        cgs_dl_scene** out)
{
    if (!out || (length && !text)) return cgs::setError(CGS_ERR_INVALID, "cgs_dl_parse: null argument");
    *out = nullptr;
    return parseInto(std::string(text ? text : "", length), name ? name : "<text>", working_dir ? working_dir : "",
                     out);
}

cgs_status cgs_dl_read(const char* path, cgs_dl_scene** out)
{
    if (!path || !out) return cgs::setError(CGS_ERR_INVALID, "cgs_dl_read: null argument");
    *out = nullptr;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return cgs::setError(CGS_ERR_INVALID, std::string("cannot open '") + path + "'");
    std::string text;
    char buf[65536];
    for (size_t got; (got = std::fread(buf, 1, sizeof buf, fp)) > 0;) text.append(buf, got);
    std::fclose(fp);
    std::string p = path;
    const size_t slash = p.find_last_of('/');
    const std::string dir = slash == std::string::npos ? "" : p.substr(0, slash);
    return parseInto(text, p, dir, out);
}

void cgs_dl_free(cgs_dl_scene* scene)
{
    if (scene) delete static_cast<DlScene*>(scene->storage);
}
