// Host scene loader, BVH build and record flattening for the product library.
// Grammar and semantics follow src/scene.rs:247-571 of the reference; the
// BVH follows src/bvh.rs:52-91,166-309 exactly (same bins, SAH sweep,
// partition order and child allocation), because the device traversal must
// visit objects in the reference's order to reproduce its distances bit for bit.
#include "host_scene.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace ifx {

// ------------------------------------------------------------------ limits
bool validate_cfg(const HCfg& c, std::string& err) {
    std::ostringstream os;
    if (c.width == 0 || c.height == 0) {
        err = "width and height must be at least 1";
        return false;
    }
    if (c.width > 16384u || c.height > 16384u) {
        os << "image is " << c.width << "x" << c.height << ", which exceeds the 16384 px limit on either side";
        err = os.str();
        return false;
    }
    uint64_t px = (uint64_t)c.width * c.height;
    if (px > 40000000ull) {
        os << "image is " << px << " pixels (" << c.width << "x" << c.height << "); the limit is 40000000";
        err = os.str();
        return false;
    }
    if (c.surfels == 0) {
        err = "surfels must be at least 1";
        return false;
    }
    if (c.surfels > 2000000ull) {
        os << "surfels is " << c.surfels << "; the limit is 2000000";
        err = os.str();
        return false;
    }
    if (c.bounces == 0) {
        err = "bounces must be at least 1";
        return false;
    }
    if (c.bounces > 4096u) {
        os << "bounces is " << c.bounces << "; the limit is 4096";
        err = os.str();
        return false;
    }
    if (!std::isfinite(c.exposure) || c.exposure <= 0.0f) {
        os << "exposure must be a positive finite number, got " << c.exposure;
        err = os.str();
        return false;
    }
    return true;
}

// ------------------------------------------------------------------ lexer
namespace {

// Token stream of a `.rad` file: punctuation {}[],: as single tokens,
// whitespace-separated words, `#` comments to end of line, and quoted
// strings kept with a leading '"' marker (src/scene.rs:247-300).
std::vector<std::string> lex(const std::string& s) {
    std::vector<std::string> toks;
    std::string word;
    auto flush = [&]() {
        if (!word.empty()) toks.push_back(word), word.clear();
    };
    size_t i = 0, n = s.size();
    while (i < n) {
        char c = s[i++];
        if (c == '"') {
            flush();
            std::string str("\"");
            bool closed = false;
            while (i < n) {
                char d = s[i++];
                if (d == '"') {
                    closed = true;
                    break;
                }
                str.push_back(d);
            }
            if (closed)
                toks.push_back(str);
            else
                word = str.substr(1);  // unterminated: the text survives as a bare word
            continue;
        }
        if (c == '#') {
            while (i < n && s[i] != '\n') ++i;
            if (i < n) ++i;
            continue;
        }
        if (std::strchr("{}[],:", c) && c != 0) {
            flush();
            toks.push_back(std::string(1, c));
            continue;
        }
        if (c == ' ' || (c >= '\t' && c <= '\r')) {
            flush();
            continue;
        }
        word.push_back(c);
    }
    flush();
    return toks;
}

struct Reader {
    std::vector<std::string> t;
    size_t pos = 0;
    std::string err;
    bool ok() const { return err.empty(); }
    bool at(const char* s) const { return pos < t.size() && t[pos] == s; }
    bool more() const { return pos < t.size(); }
    std::string take() {
        if (!ok()) return std::string();
        if (pos >= t.size()) {
            err = "unexpected end of file";
            return std::string();
        }
        return t[pos++];
    }
    void want(const char* s) {
        std::string g = take();
        if (ok() && g != s) err = std::string("expected `") + s + "`, found `" + g + "`";
    }
    void opt_comma() {
        if (ok() && at(",")) ++pos;
    }
    // Rust's f32 grammar: optional sign, decimal / inf / nan, whole token.
    float number() {
        std::string g = take();
        if (!ok()) return 0.0f;
        const char* b = g.c_str();
        const char* q = b + ((b[0] == '+' || b[0] == '-') ? 1 : 0);
        bool bad = g.empty() || (q[0] == '0' && (q[1] == 'x' || q[1] == 'X')) || g.find('(') != std::string::npos ||
                   g.find_first_of(" \t\n\r\v\f") != std::string::npos;
        char* end = nullptr;
        float v = bad ? 0.0f : std::strtof(b, &end);
        if (bad || end == b || *end != '\0') {
            err = "expected number, found `" + g + "`";
            return 0.0f;
        }
        return v;
    }
    f3 vec() {
        want("[");
        float x = number();
        opt_comma();
        float y = number();
        opt_comma();
        float z = number();
        opt_comma();
        want("]");
        return mk(x, y, z);
    }
    std::string name() {
        std::string g = take();
        size_t k = g.find_first_not_of('"');
        return k == std::string::npos ? std::string() : g.substr(k);
    }
    // Block body: calls `on_key(key)` for every `key:` until the closing brace.
    void body(const std::function<void(const std::string&)>& on_key) {
        want("{");
        while (ok()) {
            if (at("}")) {
                ++pos;
                return;
            }
            std::string k = take();
            want(":");
            if (!ok()) return;
            on_key(k);
            opt_comma();
        }
    }
};

float sat_from(float v) { return v; }

}  // namespace

// ------------------------------------------------------------------ BVH
namespace {

struct BuildBox {
    f3 lo, hi, c;
};
struct BuildNode {
    f3 lo, hi;
    uint32_t start = 0, count = 0, right = 0;
};

struct BvhBuilder {
    const std::vector<BuildBox>& box;
    std::vector<uint32_t>& ord;
    std::vector<BuildNode>& out;
    static constexpr int kBins = 12;
    static constexpr size_t kLeaf = 4;

    static float area(f3 lo, f3 hi) {
        if (lo.x > hi.x) return 0.0f;
        f3 d = hi - lo;
        return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
    }
    static size_t bin_of(float c, float c0, float scale) {
        float f = (c - c0) * scale;
        uint64_t b = !(f > 0.0f) ? 0ull : (f >= 1.8446744e19f ? ~0ull : (uint64_t)f);
        return b < (uint64_t)(kBins - 1) ? (size_t)b : (size_t)(kBins - 1);
    }

    void node(size_t at, size_t first, size_t n) {
        f3 lo = spl(kFltMax), hi = spl(-kFltMax), clo = spl(kFltMax), chi = spl(-kFltMax);
        for (size_t k = first; k < first + n; ++k) {
            const BuildBox& b = box[ord[k]];
            lo = min3(lo, b.lo);
            hi = max3(hi, b.hi);
            clo = min3(clo, b.c);
            chi = max3(chi, b.c);
        }
        out[at].lo = lo;
        out[at].hi = hi;
        auto make_leaf = [&]() {
            out[at].start = (uint32_t)first;
            out[at].count = (uint32_t)n;
        };
        if (n <= kLeaf) return make_leaf();
        f3 ext = chi - clo;
        int ax = (ext.x > ext.y && ext.x > ext.z) ? 0 : (ext.y > ext.z ? 1 : 2);
        float axe = at3(ext, ax);
        if (axe < 1e-9f) return make_leaf();
        float scale = (float)kBins / axe;
        float c0 = at3(clo, ax);
        f3 blo[kBins], bhi[kBins];
        size_t bn[kBins];
        for (int i = 0; i < kBins; ++i) blo[i] = spl(kFltMax), bhi[i] = spl(-kFltMax), bn[i] = 0;
        for (size_t k = first; k < first + n; ++k) {
            const BuildBox& b = box[ord[k]];
            size_t bi = bin_of(at3(b.c, ax), c0, scale);
            blo[bi] = min3(blo[bi], b.lo);
            bhi[bi] = max3(bhi[bi], b.hi);
            bn[bi]++;
        }
        // Left prefix areas and counts, then a right-to-left sweep for the
        // cheapest boundary (ties keep the rightmost boundary).
        float la[kBins];
        size_t lc[kBins];
        f3 al = spl(kFltMax), ah = spl(-kFltMax);
        size_t acc = 0;
        for (int i = 0; i < kBins; ++i) {
            al = min3(al, blo[i]);
            ah = max3(ah, bhi[i]);
            acc += bn[i];
            la[i] = area(al, ah);
            lc[i] = acc;
        }
        float best = kFltMax;
        int split = kBins / 2;
        f3 rl = spl(kFltMax), rh = spl(-kFltMax);
        size_t rc = 0;
        for (int i = kBins - 1; i >= 1; --i) {
            rl = min3(rl, blo[i]);
            rh = max3(rh, bhi[i]);
            rc += bn[i];
            if (lc[i - 1] == 0 || rc == 0) continue;
            float cost = la[i - 1] * (float)lc[i - 1] + area(rl, rh) * (float)rc;
            if (cost < best) best = cost, split = i;
        }
        size_t mid = 0;
        for (size_t k = 0; k < n; ++k) {
            if ((int)bin_of(at3(box[ord[first + k]].c, ax), c0, scale) < split) {
                std::swap(ord[first + k], ord[first + mid]);
                ++mid;
            }
        }
        if (mid == 0 || mid == n) mid = n / 2;
        size_t left = out.size();
        out.emplace_back();
        node(left, first, mid);
        size_t right = out.size();
        out.emplace_back();
        out[at].count = 0;
        out[at].right = (uint32_t)right;
        node(right, first + mid, n - mid);
    }
};

bool prim_bounds(const HPrim& p, f3& lo, f3& hi) {
    switch (p.kind) {
        case K_SPHERE:
            lo = p.c - spl(p.r), hi = p.c + spl(p.r);
            return true;
        case K_ROUNDBOX:
            lo = p.c - p.b, hi = p.c + p.b;
            return true;
        case K_PLANE:
            return false;
        case K_CAPSULE:
        case K_CYLINDER:
            lo = min3(p.a, p.b) - spl(p.r), hi = max3(p.a, p.b) + spl(p.r);
            return true;
        case K_TORUS: {
            f3 e = mk(p.major + p.minor, p.minor, p.major + p.minor);
            lo = p.c - e, hi = p.c + e;
            return true;
        }
        default: {
            float r = fmx(p.ra, p.rb);
            lo = min3(p.a, p.b) - spl(r), hi = max3(p.a, p.b) + spl(r);
            return true;
        }
    }
}

void prim_rec(const HPrim& p, float4* r) {
    float4 g0 = make_float4(0, 0, 0, 0), g1 = g0, g2 = g0;
    switch (p.kind) {
        case K_SPHERE:
            g0 = st4(p.c, p.r);
            break;
        case K_ROUNDBOX:
            g0 = st4(p.c, p.r);
            g1 = st4(p.b - spl(p.r));
            break;
        case K_PLANE:
            g0 = st4(p.n, p.h);
            break;
        case K_CAPSULE:
        case K_CYLINDER: {
            f3 ba = p.b - p.a;
            g0 = st4(p.a, p.r);
            g1 = st4(ba, dot3(ba, ba));
            break;
        }
        case K_TORUS:
            g0 = st4(p.c, p.major);
            g2.x = p.minor;
            break;
        default: {
            f3 ba = p.b - p.a;
            g0 = st4(p.a, p.ra);
            g1 = st4(ba, dot3(ba, ba));
            g2.x = p.rb;
            g2.y = p.rb - p.ra;
        }
    }
    g2.z = bitsf(p.kind);
    r[0] = g0, r[1] = g1, r[2] = g2;
    r[3] = make_float4(0, 0, 0, 0);
}

}  // namespace

namespace {
void flatten_scene(HostScene& s, const std::vector<BuildNode>& nodes, const f3* bounds);
}  // namespace

void finalize_scene(HostScene& s, const f3* bounds) {
    // BVH over bounded objects, boxes grown by the blend radius (src/bvh.rs:62-72).
    std::vector<BuildBox> boxes;
    for (const HObject& o : s.objects) {
        f3 lo, hi;
        prim_bounds(o.prim, lo, hi);
        lo = lo - spl(o.blend);
        hi = hi + spl(o.blend);
        boxes.push_back(BuildBox{lo, hi, (lo + hi) * 0.5f});
    }
    std::vector<BuildNode> nodes;
    s.order.resize(s.objects.size());
    for (size_t i = 0; i < s.order.size(); ++i) s.order[i] = (uint32_t)i;
    if (!s.objects.empty()) {
        nodes.reserve(2 * s.objects.size());
        nodes.emplace_back();
        BvhBuilder b{boxes, s.order, nodes};
        b.node(0, 0, s.objects.size());
    }
    flatten_scene(s, nodes, bounds);
}

void finalize_scene_given(HostScene& s, const std::vector<HBvhNode>& given, const std::vector<uint32_t>& order,
                          f3 lo, f3 hi) {
    if (order.size() != s.objects.size()) throw std::runtime_error("BVH order must list every object once");
    std::vector<uint32_t> seen(order.size(), 0u);
    for (uint32_t o : order)
        if (o >= order.size() || seen[o]++) throw std::runtime_error("BVH order must list every object once");
    if (given.empty() != s.objects.empty()) throw std::runtime_error("a scene with objects needs BVH nodes");
    std::vector<BuildNode> nodes(given.size());
    for (size_t i = 0; i < given.size(); ++i) {
        nodes[i].lo = given[i].lo;
        nodes[i].hi = given[i].hi;
        nodes[i].start = given[i].start;
        nodes[i].count = given[i].count;
        nodes[i].right = given[i].right;
    }
    s.order = order;
    const f3 b[2] = {lo, hi};
    flatten_scene(s, nodes, b);
}

namespace {
// Device records from the objects, the BVH (nodes over s.order) and the
// bounds (world_bounds when `bounds` is null).  Throws std::runtime_error
// on a BVH the traversals cannot walk.
void flatten_scene(HostScene& s, const std::vector<BuildNode>& nodes, const f3* bounds) {
    // Subtraction primitives, shared by objects and planes.
    s.sub_rec.clear();
    auto object_rec = [&](const HObject& o, float4* r) {
        prim_rec(o.prim, r);
        uint32_t sb = (uint32_t)(s.sub_rec.size() / 4);
        for (const HPrim& q : o.subs) {
            float4 t[4];
            prim_rec(q, t);
            s.sub_rec.insert(s.sub_rec.end(), t, t + 4);
        }
        r[3] = make_float4(bitsf(o.mat), bitsf(sb), bitsf((uint32_t)o.subs.size()), bitsf(o.blend > 0.0f ? 1u : 0u));
    };

    // The device traversals rely on this layout: every slot lies in exactly
    // one non-empty leaf, and each interior node's children (node + 1 and
    // `right`) lie after it, so a traversal always ends.
    {
        const size_t n_obj = s.objects.size();
        std::vector<uint32_t> cover(n_obj, 0u);
        for (size_t ni = 0; ni < nodes.size(); ++ni) {
            const BuildNode& n = nodes[ni];
            if (n.count > 0) {
                for (uint64_t k = n.start; k < (uint64_t)n.start + n.count; ++k)
                    if (k >= n_obj || cover[k]++) throw std::runtime_error("BVH leaf ranges overlap or overrun");
            } else if (!(ni + 1 < nodes.size() && n.right > ni + 1 && n.right < nodes.size())) {
                throw std::runtime_error("BVH interior node has an out-of-order child");
            }
        }
        for (size_t k = 0; k < n_obj; ++k)
            if (cover[k] != 1u) throw std::runtime_error("BVH leaves do not cover every object");
    }
    s.node_rec.resize(2 * nodes.size());
    for (size_t i = 0; i < nodes.size(); ++i) {
        const BuildNode& n = nodes[i];
        s.node_rec[2 * i] = st4(n.lo, bitsf(n.count > 0 ? n.start : n.right));
        s.node_rec[2 * i + 1] = st4(n.hi, bitsf(n.count));
    }
    // Deepest traversal stack: an interior node at depth d is popped with at
    // most d pending siblings beneath it and pushes two children.
    s.max_stack = nodes.empty() ? 0 : 1;
    {
        std::vector<std::pair<uint32_t, uint32_t>> todo{{0u, 0u}};
        while (!nodes.empty() && !todo.empty()) {
            auto [ni, d] = todo.back();
            todo.pop_back();
            if (nodes[ni].count == 0) {
                s.max_stack = std::max<uint32_t>(s.max_stack, d + 2);
                todo.push_back({ni + 1, d + 1});
                todo.push_back({nodes[ni].right, d + 1});
            }
        }
    }

    // Objects in leaf order so a leaf's range indexes the records directly.
    s.obj_rec.assign(4 * s.objects.size(), make_float4(0, 0, 0, 0));
    // lin[i] = slot of object i (linear scan order); lin[n + k] = leaf node of slot k.
    const size_t n_obj = s.objects.size();
    s.lin.assign(2 * n_obj, 0);
    for (size_t slot = 0; slot < s.order.size(); ++slot) {
        object_rec(s.objects[s.order[slot]], &s.obj_rec[4 * slot]);
        s.lin[s.order[slot]] = (uint32_t)slot;
    }
    for (size_t ni = 0; ni < nodes.size(); ++ni)
        for (uint32_t k = nodes[ni].start; nodes[ni].count > 0 && k < nodes[ni].start + nodes[ni].count; ++k)
            s.lin[n_obj + k] = (uint32_t)ni;
    // Each object's box as the BVH build grows it (src/bvh.rs:62-72), with
    // its leaf node's index in lo.w.
    s.obox_rec.assign(2 * s.objects.size(), make_float4(0, 0, 0, 0));
    for (size_t slot = 0; slot < s.order.size(); ++slot) {
        const HObject& o = s.objects[s.order[slot]];
        f3 lo, hi;
        prim_bounds(o.prim, lo, hi);
        s.obox_rec[2 * slot] = st4(lo - spl(o.blend), bitsf(s.lin[n_obj + slot]));
        s.obox_rec[2 * slot + 1] = st4(hi + spl(o.blend));
    }

    s.plane_rec.assign(4 * s.planes.size(), make_float4(0, 0, 0, 0));
    for (size_t i = 0; i < s.planes.size(); ++i) object_rec(s.planes[i], &s.plane_rec[4 * i]);
    s.blend_k = 0.0f;
    for (const HObject& o : s.objects) s.blend_k = fmx(s.blend_k, o.blend);

    // World bounds (src/scene.rs:573-601; EXT emitters included).
    f3 lo = spl(kFltMax), hi = spl(-kFltMax);
    if (bounds) {
        s.lo = bounds[0], s.hi = bounds[1];
    } else {
    for (const HObject& o : s.objects) {
        f3 a, b;
        if (prim_bounds(o.prim, a, b)) lo = min3(lo, a), hi = max3(hi, b);
    }
    for (const HQuad& q : s.quads)
        for (int k = 0; k < 4; ++k) lo = min3(lo, q.v[k]), hi = max3(hi, q.v[k]);
    for (const HSphereLight& q : s.spheres) lo = min3(lo, q.c - spl(q.r)), hi = max3(hi, q.c + spl(q.r));
    for (const HPointLight& q : s.points) lo = min3(lo, q.pos), hi = max3(hi, q.pos);
    if (lo.x > hi.x) {
        s.lo = spl(-5.0f), s.hi = spl(5.0f);
    } else {
        float dg = len3(hi - lo);
        float pad = s.planes.empty() ? dg * 0.02f : dg * 0.25f;
        s.lo = lo - spl(pad), s.hi = hi + spl(pad);
    }
    }

    // Emitters and materials with the reference's per-call derived terms.
    s.quad_rec.clear();
    for (const HQuad& q : s.quads) {
        f3 e1 = q.v[1] - q.v[0], e3 = q.v[3] - q.v[0];
        float a = len3(cross3(e1, e3));
        float b = len3(cross3(q.v[3] - q.v[2], q.v[1] - q.v[2]));
        float area = 0.5f * (a + b);
        f3 ctr = (((q.v[0] + q.v[1]) + q.v[2]) + q.v[3]) * 0.25f;
        float radius = fmx(sqrtf(area / kPi), 1e-4f);
        for (int k = 0; k < 4; ++k) s.quad_rec.push_back(st4(q.v[k]));
        s.quad_rec.push_back(st4(q.emit, area));
        s.quad_rec.push_back(st4(ctr, radius));
        s.quad_rec.push_back(st4(nrm3(cross3(e1, e3))));
    }
    s.sph_rec.clear();
    for (const HSphereLight& q : s.spheres) s.sph_rec.push_back(st4(q.c, q.r)), s.sph_rec.push_back(st4(q.emit));
    s.pnt_rec.clear();
    for (const HPointLight& q : s.points) s.pnt_rec.push_back(st4(q.pos)), s.pnt_rec.push_back(st4(q.emit));
    s.mat_rec.clear();
    for (const HMaterial& m : s.mats) {
        float4 r[4];
        material_records(m, r);
        s.mat_rec.insert(s.mat_rec.end(), r, r + 4);
    }
}

}  // namespace

uint64_t scene_digest(const HostScene& s) {
    uint64_t h = 1469598103934665603ull;  // FNV-1a, 64 bit
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* b = (const unsigned char*)p;
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    };
    auto vec = [&](const auto& v) {
        const uint64_t n = v.size();
        mix(&n, sizeof(n));
        if (n) mix(v.data(), n * sizeof(v[0]));
    };
    vec(s.obj_rec), vec(s.plane_rec), vec(s.sub_rec), vec(s.node_rec), vec(s.order), vec(s.lin);
    vec(s.quad_rec), vec(s.sph_rec), vec(s.pnt_rec), vec(s.mat_rec);
    const uint32_t counts[6] = {(uint32_t)s.objects.size(), (uint32_t)s.planes.size(), (uint32_t)s.quads.size(),
                                (uint32_t)s.spheres.size(), (uint32_t)s.points.size(), s.max_stack};
    mix(counts, sizeof(counts));
    const float f[18] = {s.lo.x, s.lo.y, s.lo.z, s.hi.x, s.hi.y, s.hi.z, s.blend_k, s.cam.pos.x, s.cam.pos.y,
                         s.cam.pos.z, s.cam.look.x, s.cam.look.y, s.cam.look.z, s.cam.up.x, s.cam.up.y, s.cam.up.z,
                         s.cam.fov, s.cfg.exposure};
    mix(f, sizeof(f));
    const uint64_t c[4] = {s.cfg.width, s.cfg.height, s.cfg.surfels, s.cfg.bounces};
    mix(c, sizeof(c));
    return h;
}

void material_records(const HMaterial& m, float4 out[4]) {
    out[0] = st4(m.albedo, m.roughness);
    out[1] = st4(m.emissive, m.metallic);
    out[2] = st4(lerp3(spl(0.04f), m.albedo, m.metallic));  // f0 (src/shade.rs:36-38)
    out[3] = st4(m.albedo * (1.0f - m.metallic));          // kd (src/shade.rs:41-43)
}

SceneView HostScene::host_view() const {
    SceneView v;
    v.obj = obj_rec.data();
    v.planes = plane_rec.data();
    v.subs = sub_rec.data();
    v.nodes = node_rec.data();
    v.lin = lin.data();
    v.n_obj = (uint32_t)objects.size();
    v.n_planes = (uint32_t)planes.size();
    v.n_nodes = (uint32_t)(node_rec.size() / 2);
    v.use_bvh = 1;
    v.n_subs = (uint32_t)(sub_rec.size() / 4);
    v.blend_k = blend_k;
    v.keep_global();
    v.g_obox = obox_rec.data();
    return v;
}

LightsView HostScene::host_lights() const {
    LightsView l;
    l.quad = quad_rec.data();
    l.sph = sph_rec.data();
    l.pnt = pnt_rec.data();
    l.mat = mat_rec.data();
    l.n_quad = (uint32_t)quads.size();
    l.n_sph = (uint32_t)spheres.size();
    l.n_pnt = (uint32_t)points.size();
    l.n_mat = (uint32_t)mats.size();
    return l;
}

// ------------------------------------------------------------------ parser
bool load_rad(const std::string& src, HostScene& s, std::string& err) {
    s = HostScene();
    s.mats.push_back(HMaterial());  // material 0: the implicit default
    std::unordered_map<std::string, uint32_t> names;
    Reader R;
    R.t = lex(src);

    static const char* kShapes[] = {"sphere", "box", "plane", "capsule", "torus", "cylinder", "cone"};
    while (R.ok() && R.more()) {
        std::string blk = R.take();
        if (blk == "camera") {
            R.body([&](const std::string& k) {
                if (k == "pos") s.cam.pos = R.vec();
                else if (k == "look") s.cam.look = R.vec();
                else if (k == "up") s.cam.up = R.vec();
                else if (k == "fov") s.cam.fov = R.number();
                else R.err = "unknown camera key `" + k + "`";
            });
        } else if (blk == "render") {
            R.body([&](const std::string& k) {
                if (k == "width") s.cfg.width = sat_u32(R.number());
                else if (k == "height") s.cfg.height = sat_u32(R.number());
                else if (k == "surfels") {
                    float v = R.number();
                    s.cfg.surfels = !(v > 0.0f) ? 0ull : (v >= 1.8446744e19f ? ~0ull : (uint64_t)v);
                } else if (k == "bounces") s.cfg.bounces = sat_u32(R.number());
                else if (k == "exposure") s.cfg.exposure = R.number();
                else R.err = "unknown render key `" + k + "`";
            });
        } else if (blk == "material") {
            std::string nm = R.name();
            HMaterial m;
            R.body([&](const std::string& k) {
                if (k == "albedo") m.albedo = R.vec();
                else if (k == "roughness") m.roughness = clmp(R.number(), 0.02f, 1.0f);
                else if (k == "metallic") m.metallic = clmp(R.number(), 0.0f, 1.0f);
                else if (k == "emissive") m.emissive = R.vec();
                else R.err = "unknown material key `" + k + "`";
            });
            names[nm] = (uint32_t)s.mats.size();
            s.mats.push_back(m);
        } else if (blk == "light") {
            HQuad q;
            for (int k = 0; k < 4; ++k) q.v[k] = spl(0.0f);
            q.emit = spl(10.0f);
            R.body([&](const std::string& k) {
                if (k == "verts") {
                    R.want("[");
                    for (int j = 0; j < 4 && R.ok(); ++j) q.v[j] = R.vec(), R.opt_comma();
                    R.want("]");
                } else if (k == "emit") {
                    q.emit = R.vec();
                } else {
                    R.err = "unknown light key `" + k + "`";
                }
            });
            s.quads.push_back(q);
        } else if (blk == "sphere_light") {  // EXT
            HSphereLight q{spl(0.0f), 0.25f, spl(10.0f)};
            R.body([&](const std::string& k) {
                if (k == "center") q.c = R.vec();
                else if (k == "radius") q.r = R.number();
                else if (k == "emit") q.emit = R.vec();
                else R.err = "unknown sphere_light key `" + k + "`";
            });
            if (R.ok() && !(q.r > 0.0f)) R.err = "sphere_light radius must be positive";
            s.spheres.push_back(q);
        } else if (blk == "point_light") {  // EXT
            HPointLight q{spl(0.0f), spl(10.0f)};
            R.body([&](const std::string& k) {
                if (k == "pos") q.pos = R.vec();
                else if (k == "emit") q.emit = R.vec();
                else R.err = "unknown point_light key `" + k + "`";
            });
            s.points.push_back(q);
        } else {
            int shape = -1;
            for (int i = 0; i < 7; ++i)
                if (blk == kShapes[i]) shape = i;
            if (shape < 0) {
                R.err = "unknown block `" + blk + "`";
                break;
            }
            HObject o;
            f3 center = spl(0.0f), size = spl(1.0f), normal = mk(0, 1, 0), a = spl(0.0f), b = mk(0, 1, 0);
            float radius = 1.0f, round = 0.0f, height = 0.0f, major = 1.0f, minor = 0.25f, rtop = 0.0f, blend = 0.0f;
            R.body([&](const std::string& k) {
                if (k == "center") center = R.vec();
                else if (k == "radius") radius = R.number();
                else if (k == "size") size = R.vec();
                else if (k == "round") round = R.number();
                else if (k == "normal") normal = nrm3(R.vec());
                else if (k == "height") height = R.number();
                else if (k == "a") a = R.vec();
                else if (k == "b") b = R.vec();
                else if (k == "major") major = R.number();
                else if (k == "minor") minor = R.number();
                else if (k == "radius_top") rtop = R.number();
                else if (k == "mat") {
                    std::string nm = R.name();
                    auto it = names.find(nm);
                    if (R.ok() && it == names.end()) R.err = "unknown material `" + nm + "`";
                    else if (R.ok()) o.mat = it->second;
                } else if (k == "subtract_sphere") {
                    R.want("[");
                    HPrim q;
                    q.kind = K_SPHERE;
                    q.c = R.vec();
                    R.opt_comma();
                    q.r = R.number();
                    R.opt_comma();
                    R.want("]");
                    o.subs.push_back(q);
                } else if (k == "subtract_cylinder") {  // EXT
                    R.want("[");
                    HPrim q;
                    q.kind = K_CYLINDER;
                    q.a = R.vec();
                    R.opt_comma();
                    q.b = R.vec();
                    R.opt_comma();
                    q.r = R.number();
                    R.opt_comma();
                    R.want("]");
                    o.subs.push_back(q);
                } else if (k == "blend") {  // EXT
                    blend = fmx(R.number(), 0.0f);
                } else {
                    R.err = "unknown shape key `" + k + "`";
                }
            });
            HPrim& p = o.prim;
            switch (shape) {
                case 0: p.kind = K_SPHERE, p.c = center, p.r = radius; break;
                case 1: p.kind = K_ROUNDBOX, p.c = center, p.b = size * 0.5f, p.r = fmn(round, mine(size) * 0.499f); break;
                case 2: p.kind = K_PLANE, p.n = normal, p.h = height; break;
                case 3: p.kind = K_CAPSULE, p.a = a, p.b = b, p.r = radius; break;
                case 4: p.kind = K_TORUS, p.c = center, p.major = major, p.minor = minor; break;
                case 5: p.kind = K_CYLINDER, p.a = a, p.b = b, p.r = radius; break;
                default: p.kind = K_CONE, p.a = a, p.b = b, p.ra = radius, p.rb = rtop;
            }
            if (p.kind == K_PLANE) {
                o.blend = 0.0f;
                s.planes.push_back(o);
            } else {
                o.blend = blend;
                s.objects.push_back(o);
            }
        }
    }
    if (!R.ok()) {
        err = R.err;
        return false;
    }
    if (s.quads.empty() && s.spheres.empty() && s.points.empty()) {
        err = "scene has no `light` block";
        return false;
    }
    if (!validate_cfg(s.cfg, err)) return false;
    try {
        finalize_scene(s);
    } catch (const std::runtime_error& e) {
        err = e.what();
        return false;
    }
    return true;
}

}  // namespace ifx
