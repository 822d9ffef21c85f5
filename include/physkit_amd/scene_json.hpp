// Scene files for the C++ facade: the JSON schema of PhysKit's demo and test helper
// (tests/helper/src/graphics.cpp:20-135; the files under tests/configs/), read into a
// world_desc and one object_desc per object, which populate() creates in a
// physkit_amd::world in file order (arena slots 0, 1, ...), as the reference's reader
// creates them in a physkit::world.
//
// Fields, defaults and conversions are the reference reader's:
//   * world: gravity (default 0, -9.81, 0), solver_iterations (10), title; camera,
//     window, drag and colour fields are display settings and are ignored;
//   * object: pos (required), vel (0), orientation [w, x, y, z] (identity),
//     angular_velocity (0), inertia_tensor as the diagonal [Ixx, Iyy, Izz] (1, 1, 1),
//     type "stat" | "dynam" (dynam; a dynamic object needs a mass), restitution (0.5),
//     friction (0.5) and a mesh (default a box of half extents 0.5):
//       {"name": "box", "half_extents": [x, y, z]}
//       {"name": "pyramid", "height": h, "base_size": b}
//       {"name": "sphere", "radius": r, "stacks": 16, "sectors": 32}
//     Mesh sizes are 32-bit floats (graphics.cpp:24-41), widened to double.
//   * The reference passes (height, base_size) to mesh::pyramid(base_half, height)
//     (graphics.cpp:104-106 against mesh.h:236): a file's "height" is the pyramid's base
//     half-width and its "base_size" the apex height. So does this reader.
//   * Identical mesh descriptions share one mesh (graphics.cpp:91-110).
// Errors throw std::runtime_error (unreadable file, malformed JSON) or
// std::invalid_argument (a field the schema does not allow).
#ifndef PHYSKIT_AMD_SCENE_JSON_HPP
#define PHYSKIT_AMD_SCENE_JSON_HPP

#include "world.hpp"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace physkit_amd
{
namespace scene_json
{

// ---- a small JSON reader: null, booleans, numbers, strings, arrays, objects ----
struct value
{
    enum kind_t
    {
        null_v,
        bool_v,
        number_v,
        string_v,
        array_v,
        object_v
    } kind = null_v;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<value> arr;
    std::vector<std::pair<std::string, value>> obj; // file order

    const value *find(const std::string &k) const
    {
        for (const auto &kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

class parser
{
public:
    explicit parser(const std::string &text) : M_s(text) {}
    value parse_document()
    {
        value v = parse_value();
        skip_ws();
        if (M_i != M_s.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string &M_s;
    size_t M_i = 0;

    [[noreturn]] void fail(const char *what) const
    { throw std::runtime_error(std::string("scene file: JSON ") + what + " at offset " + std::to_string(M_i)); }
    void skip_ws()
    {
        while (M_i < M_s.size() && std::isspace(static_cast<unsigned char>(M_s[M_i]))) ++M_i;
    }
    bool eat(char c)
    {
        skip_ws();
        if (M_i < M_s.size() && M_s[M_i] == c)
        {
            ++M_i;
            return true;
        }
        return false;
    }
    void expect(char c)
    {
        if (!eat(c)) fail((std::string("expected '") + c + "'").c_str());
    }
    bool word(const char *w)
    {
        const size_t n = std::char_traits<char>::length(w);
        if (M_s.compare(M_i, n, w) == 0)
        {
            M_i += n;
            return true;
        }
        return false;
    }
    std::string parse_string()
    {
        expect('"');
        std::string out;
        while (M_i < M_s.size() && M_s[M_i] != '"')
        {
            char c = M_s[M_i++];
            if (c == '\\')
            {
                if (M_i >= M_s.size()) fail("unterminated escape");
                char e = M_s[M_i++];
                switch (e)
                {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u':
                {
                    if (M_i + 4 > M_s.size()) fail("short \\u escape");
                    unsigned cp = static_cast<unsigned>(std::strtoul(M_s.substr(M_i, 4).c_str(), nullptr, 16));
                    M_i += 4;
                    if (cp < 0x80)
                        out += static_cast<char>(cp);
                    else if (cp < 0x800)
                    {
                        out += static_cast<char>(0xC0 | (cp >> 6));
                        out += static_cast<char>(0x80 | (cp & 0x3F));
                    }
                    else
                    {
                        out += static_cast<char>(0xE0 | (cp >> 12));
                        out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
                        out += static_cast<char>(0x80 | (cp & 0x3F));
                    }
                    break;
                }
                default: fail("unknown escape");
                }
            }
            else
                out += c;
        }
        if (M_i >= M_s.size()) fail("unterminated string");
        ++M_i;
        return out;
    }
    value parse_value()
    {
        skip_ws();
        if (M_i >= M_s.size()) fail("unexpected end");
        value v;
        const char c = M_s[M_i];
        if (c == '{')
        {
            ++M_i;
            v.kind = value::object_v;
            if (eat('}')) return v;
            do
            {
                skip_ws();
                std::string k = parse_string();
                expect(':');
                v.obj.emplace_back(std::move(k), parse_value());
            } while (eat(','));
            expect('}');
        }
        else if (c == '[')
        {
            ++M_i;
            v.kind = value::array_v;
            if (eat(']')) return v;
            do
                v.arr.push_back(parse_value());
            while (eat(','));
            expect(']');
        }
        else if (c == '"')
        {
            v.kind = value::string_v;
            v.str = parse_string();
        }
        else if (word("true"))
        {
            v.kind = value::bool_v;
            v.b = true;
        }
        else if (word("false"))
            v.kind = value::bool_v;
        else if (word("null"))
            v.kind = value::null_v;
        else
        {
            // a JSON number, converted by strtod (correctly rounded to double)
            const char *start = M_s.c_str() + M_i;
            char *end = nullptr;
            v.num = std::strtod(start, &end);
            if (end == start) fail("expected a value");
            v.kind = value::number_v;
            M_i += static_cast<size_t>(end - start);
        }
        return v;
    }
};

inline value parse(const std::string &text) { return parser(text).parse_document(); }

// ---- the scene schema ----
struct scene_file
{
    world_desc desc;
    std::string title;
    std::vector<object_desc> objects; // in file order: arena slots 0, 1, ...
};

namespace detail
{
[[noreturn]] inline void bad(size_t i, const std::string &what)
{ throw std::invalid_argument("scene file: object " + std::to_string(i) + ": " + what); }
inline double num(const value &v, const std::string &what)
{
    if (v.kind != value::number_v) throw std::invalid_argument("scene file: " + what + " must be a number");
    return v.num;
}
inline std::vector<double> nums(const value *v, std::vector<double> dflt, size_t n, const std::string &what)
{
    if (!v) return dflt;
    if (v->kind != value::array_v || v->arr.size() != n)
        throw std::invalid_argument("scene file: " + what + " must be an array of " + std::to_string(n) + " numbers");
    std::vector<double> out;
    for (const value &e : v->arr) out.push_back(num(e, what));
    return out;
}
// mesh sizes are single-precision in the reference's reader (graphics.cpp:24-41)
inline double f32(double x) { return static_cast<double>(static_cast<float>(x)); }
} // namespace detail

// Parses scene-file text (see the header comment for the schema).
inline scene_file read_scene_text(const std::string &text)
{
    using namespace detail;
    const value root = parse(text);
    if (root.kind != value::object_v) throw std::invalid_argument("scene file: the top level must be an object");
    scene_file out;
    const std::vector<double> g = nums(root.find("gravity"), {0.0, -9.81, 0.0}, 3, "gravity");
    out.desc.with_gravity({g[0], g[1], g[2]});
    if (const value *it = root.find("solver_iterations"))
        out.desc.with_solver_iterations(static_cast<uint32_t>(num(*it, "solver_iterations")));
    if (const value *t = root.find("title"))
        if (t->kind == value::string_v) out.title = t->str;
    const value *objs = root.find("objects");
    if (!objs) return out;
    if (objs->kind != value::array_v) throw std::invalid_argument("scene file: objects must be an array");
    std::map<std::string, std::shared_ptr<const mesh>> shared; // identical descriptions, one mesh
    for (size_t i = 0; i < objs->arr.size(); ++i)
    {
        const value &o = objs->arr[i];
        if (o.kind != value::object_v) bad(i, "not an object");
        std::string type = "dynam";
        if (const value *t = o.find("type"))
        {
            if (t->kind != value::string_v) bad(i, "type must be a string");
            type = t->str;
        }
        if (type != "stat" && type != "dynam") bad(i, "type must be 'stat' or 'dynam'");
        object_desc d = type == "stat" ? object_desc::stat() : object_desc::dynam();
        const value *pos = o.find("pos");
        if (!pos) bad(i, "pos is required");
        const std::vector<double> p = nums(pos, {}, 3, "pos");
        d.with_pos({p[0], p[1], p[2]});
        const std::vector<double> v = nums(o.find("vel"), {0.0, 0.0, 0.0}, 3, "vel");
        d.with_vel({v[0], v[1], v[2]});
        const std::vector<double> q = nums(o.find("orientation"), {1.0, 0.0, 0.0, 0.0}, 4, "orientation");
        d.with_orientation({q[0], q[1], q[2], q[3]});
        const std::vector<double> w = nums(o.find("angular_velocity"), {0.0, 0.0, 0.0}, 3, "angular_velocity");
        d.with_ang_vel({w[0], w[1], w[2]});
        const std::vector<double> I = nums(o.find("inertia_tensor"), {1.0, 1.0, 1.0}, 3, "inertia_tensor");
        d.with_inertia_tensor({I[0], 0, 0, 0, I[1], 0, 0, 0, I[2]});
        const value *m = o.find("mass");
        if (!m && type == "dynam") bad(i, "a dynamic object needs a mass");
        d.with_mass(m ? num(*m, "mass") : 0.0);
        if (const value *e = o.find("restitution")) d.with_restitution(num(*e, "restitution"));
        if (const value *f = o.find("friction")) d.with_friction(num(*f, "friction"));
        // the mesh
        std::string name = "box";
        std::vector<double> h = {0.5, 0.5, 0.5};
        double height = 0, base = 0, radius = 0;
        unsigned stacks = 16, sectors = 32;
        if (const value *mv = o.find("mesh"))
        {
            if (mv->kind != value::object_v) bad(i, "mesh must be an object");
            if (const value *nm = mv->find("name")) name = nm->kind == value::string_v ? nm->str : "";
            if (name == "box")
                h = nums(mv->find("half_extents"), {0.5, 0.5, 0.5}, 3, "half_extents");
            else if (name == "pyramid")
            {
                const value *hv = mv->find("height"), *bv = mv->find("base_size");
                if (!hv || !bv) bad(i, "a pyramid needs height and base_size");
                height = num(*hv, "height");
                base = num(*bv, "base_size");
            }
            else if (name == "sphere")
            {
                const value *rv = mv->find("radius");
                if (!rv) bad(i, "a sphere needs a radius");
                radius = num(*rv, "radius");
                if (const value *s = mv->find("stacks")) stacks = static_cast<unsigned>(num(*s, "stacks"));
                if (const value *s = mv->find("sectors")) sectors = static_cast<unsigned>(num(*s, "sectors"));
                if (stacks < 2 || sectors < 3) bad(i, "a sphere needs stacks >= 2 and sectors >= 3"); // mesh.cpp:62-63
            }
            else
                bad(i, "unknown mesh name '" + name + "'");
        }
        std::ostringstream key;
        key.precision(17);
        std::shared_ptr<const mesh> mp;
        if (name == "box")
        {
            const vec3 hx{f32(h[0]), f32(h[1]), f32(h[2])};
            key << "box " << hx.x << ' ' << hx.y << ' ' << hx.z;
            auto it = shared.find(key.str());
            mp = it != shared.end() ? it->second : (shared[key.str()] = mesh::box(hx));
        }
        else if (name == "pyramid")
        {
            // (height, base_size) into mesh::pyramid(base_half, height), as graphics.cpp:104-106
            key << "pyramid " << f32(height) << ' ' << f32(base);
            auto it = shared.find(key.str());
            mp = it != shared.end() ? it->second : (shared[key.str()] = mesh::pyramid(f32(height), f32(base)));
        }
        else
        {
            key << "sphere " << f32(radius) << ' ' << stacks << ' ' << sectors;
            auto it = shared.find(key.str());
            mp = it != shared.end() ? it->second : (shared[key.str()] = mesh::sphere(f32(radius), stacks, sectors));
        }
        d.with_mesh(mp);
        out.objects.push_back(std::move(d));
    }
    return out;
}

// Reads a scene file from disk.
inline scene_file read_scene(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("scene file: cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return read_scene_text(ss.str());
}

// Creates the scene's objects in w, in file order; returns their handles.
inline std::vector<handle> populate(world &w, const scene_file &s)
{
    std::vector<handle> out;
    out.reserve(s.objects.size());
    for (const object_desc &d : s.objects) out.push_back(w.create_rigid(d));
    return out;
}

} // namespace scene_json
} // namespace physkit_amd

#endif
