// ORACLE — test infrastructure only (see ovec.h).
// Restates src/sdf.rs, src/bvh.rs and src/scene.rs of the reference.
#include "oscene.h"

#include <algorithm>
#include <cfloat>
#include <cstdlib>
#include <map>
#include <sstream>

namespace orc {

// ---------------------------------------------------------------- sdf.rs:50-125
float Prim::dist(V3 p) const {
    switch (kind) {
        case SPHERE:
            return length(p - c) - r;
        case ROUNDBOX: {
            V3 q = vabs(p - c) - (b - splat(r));
            return (length(vmax(q, splat(0.0f))) + fmin_(fmax_(q.x, fmax_(q.y, q.z)), 0.0f)) - r;
        }
        case PLANE:
            return dot(p, n) - h;
        case CAPSULE: {
            V3 pa = p - a;
            V3 ba = b - a;
            float hh = clampf(dot(pa, ba) / dot(ba, ba), 0.0f, 1.0f);
            return length(pa - ba * hh) - r;
        }
        case TORUS: {
            V3 q = p - c;
            float xz = std::sqrt(q.x * q.x + q.z * q.z) - major;
            return std::sqrt(xz * xz + q.y * q.y) - minor;
        }
        case CYLINDER: {
            V3 ba = b - a;
            V3 pa = p - a;
            float baba = dot(ba, ba);
            if (baba < 1e-12f) return length(p - a) - r;
            float paba = dot(pa, ba);
            float x = length(pa * baba - ba * paba) - r * baba;
            float y = std::fabs(paba - baba * 0.5f) - baba * 0.5f;
            float x2 = x * x;
            float y2 = y * y * baba;
            float d;
            if (fmax_(x, y) < 0.0f) {
                d = -fmin_(x2, y2);
            } else {
                d = (x > 0.0f ? x2 : 0.0f) + (y > 0.0f ? y2 : 0.0f);
            }
            float sg = std::signbit(d) ? -1.0f : 1.0f;  // f32::signum (d is never NaN here)
            return sg * std::sqrt(std::fabs(d)) / baba;
        }
        case CONE: {
            V3 ba = b - a;
            float baba = dot(ba, ba);
            if (baba < 1e-12f) return length(p - a) - fmax_(ra, rb);
            float rba = rb - ra;
            V3 pa = p - a;
            float papa = dot(pa, pa);
            float paba = dot(pa, ba) / baba;
            float x = std::sqrt(fmax_(papa - paba * paba * baba, 0.0f));
            float cax = fmax_(0.0f, x - (paba < 0.5f ? ra : rb));
            float cay = std::fabs(paba - 0.5f) - 0.5f;
            float k = rba * rba + baba;
            float f = clampf((rba * (x - ra) + paba * baba) / k, 0.0f, 1.0f);
            float cbx = x - ra - f * rba;
            float cby = paba - f;
            float s = (cbx < 0.0f && cay < 0.0f) ? -1.0f : 1.0f;
            return s * std::sqrt(fmin_(cax * cax + cay * cay * baba, cbx * cbx + cby * cby * baba));
        }
    }
    return FLT_MAX;
}

// ---------------------------------------------------------------- sdf.rs:127-148
bool Prim::bounds(V3& lo, V3& hi) const {
    switch (kind) {
        case SPHERE:
            lo = c - splat(r);
            hi = c + splat(r);
            return true;
        case ROUNDBOX:
            lo = c - b;
            hi = c + b;
            return true;
        case PLANE:
            return false;
        case CAPSULE:
        case CYLINDER:
            lo = vmin(a, b) - splat(r);
            hi = vmax(a, b) + splat(r);
            return true;
        case TORUS: {
            V3 e = v3(major + minor, minor, major + minor);
            lo = c - e;
            hi = c + e;
            return true;
        }
        case CONE: {
            float rr = fmax_(ra, rb);
            lo = vmin(a, b) - splat(rr);
            hi = vmax(a, b) + splat(rr);
            return true;
        }
    }
    return false;
}

// ---------------------------------------------------------------- bvh.rs
static const int N_BINS = 12;
static const size_t LEAF_SIZE = 4;

// bvh.rs:45-50
static inline float aabb_dist(V3 p, V3 lo, V3 hi) {
    V3 d = vmax(vmax(lo - p, p - hi), splat(0.0f));
    return length(d);
}

struct Box3 {
    V3 lo, hi, c;
};

static float sa_of(V3 lo, V3 hi) {
    if (lo.x > hi.x) return 0.0f;
    V3 d = hi - lo;
    return 2.0f * ((d.x * d.y + d.y * d.z) + d.z * d.x);
}

static inline size_t as_usize(float f) { return (size_t)as_u64(f); }

// bvh.rs:166-309
static void build_recursive(std::vector<BvhNode>& nodes, std::vector<uint32_t>& order,
                            const std::vector<Box3>& boxes, size_t node_idx, size_t start,
                            size_t count) {
    V3 lo = splat(FLT_MAX), hi = splat(-FLT_MAX), clo = splat(FLT_MAX), chi = splat(-FLT_MAX);
    for (size_t k = start; k < start + count; ++k) {
        const Box3& bx = boxes[order[k]];
        lo = vmin(lo, bx.lo);
        hi = vmax(hi, bx.hi);
        clo = vmin(clo, bx.c);
        chi = vmax(chi, bx.c);
    }
    nodes[node_idx].lo = lo;
    nodes[node_idx].hi = hi;
    if (count <= LEAF_SIZE) {
        nodes[node_idx].start = (uint32_t)start;
        nodes[node_idx].count = (uint32_t)count;
        return;
    }
    V3 ext = chi - clo;
    int axis = (ext.x > ext.y && ext.x > ext.z) ? 0 : (ext.y > ext.z ? 1 : 2);
    float axis_ext = comp(ext, axis);
    if (axis_ext < 1e-9f) {
        nodes[node_idx].start = (uint32_t)start;
        nodes[node_idx].count = (uint32_t)count;
        return;
    }
    float scale = (float)N_BINS / axis_ext;
    V3 bin_lo[N_BINS], bin_hi[N_BINS];
    size_t bin_n[N_BINS];
    for (int i = 0; i < N_BINS; ++i) {
        bin_lo[i] = splat(FLT_MAX);
        bin_hi[i] = splat(-FLT_MAX);
        bin_n[i] = 0;
    }
    for (size_t k = start; k < start + count; ++k) {
        const Box3& bx = boxes[order[k]];
        size_t bi = std::min(as_usize((comp(bx.c, axis) - comp(clo, axis)) * scale), (size_t)(N_BINS - 1));
        bin_lo[bi] = vmin(bin_lo[bi], bx.lo);
        bin_hi[bi] = vmax(bin_hi[bi], bx.hi);
        bin_n[bi] += 1;
    }
    float left_area[N_BINS];
    size_t left_cnt[N_BINS];
    V3 al = splat(FLT_MAX), ah = splat(-FLT_MAX);
    size_t ac = 0;
    for (int i = 0; i < N_BINS; ++i) {
        al = vmin(al, bin_lo[i]);
        ah = vmax(ah, bin_hi[i]);
        ac += bin_n[i];
        left_area[i] = sa_of(al, ah);
        left_cnt[i] = ac;
    }
    float best_cost = FLT_MAX;
    int best_split = N_BINS / 2;
    V3 bl = splat(FLT_MAX), bh = splat(-FLT_MAX);
    size_t bc = 0;
    for (int i = N_BINS - 1; i >= 1; --i) {
        bl = vmin(bl, bin_lo[i]);
        bh = vmax(bh, bin_hi[i]);
        bc += bin_n[i];
        size_t lc = left_cnt[i - 1];
        if (lc == 0 || bc == 0) continue;
        float cost = left_area[i - 1] * (float)lc + sa_of(bl, bh) * (float)bc;
        if (cost < best_cost) {
            best_cost = cost;
            best_split = i;
        }
    }
    size_t mid = 0;
    for (size_t i = 0; i < count; ++i) {
        const Box3& bx = boxes[order[start + i]];
        size_t bi = std::min(as_usize((comp(bx.c, axis) - comp(clo, axis)) * scale), (size_t)(N_BINS - 1));
        if ((int)bi < best_split) {
            std::swap(order[start + i], order[start + mid]);
            mid += 1;
        }
    }
    if (mid == 0 || mid == count) mid = count / 2;

    size_t left = nodes.size();
    nodes.push_back(BvhNode{});
    build_recursive(nodes, order, boxes, left, start, mid);
    size_t right = nodes.size();
    nodes.push_back(BvhNode{});
    nodes[node_idx].count = 0;
    nodes[node_idx].right = (uint32_t)right;
    build_recursive(nodes, order, boxes, right, start + mid, count - mid);
}

// bvh.rs:53-91
void Bvh::build(const std::vector<Object>& objects) {
    nodes.clear();
    order.clear();
    if (objects.empty()) {
        empty = true;
        return;
    }
    empty = false;
    std::vector<Box3> boxes;
    boxes.reserve(objects.size());
    for (const Object& o : objects) {
        V3 lo, hi;
        o.prim.bounds(lo, hi);
        V3 e = splat(o.blend);
        lo = lo - e;
        hi = hi + e;
        boxes.push_back(Box3{lo, hi, (lo + hi) * 0.5f});
    }
    order.resize(objects.size());
    for (size_t i = 0; i < objects.size(); ++i) order[i] = (uint32_t)i;
    nodes.reserve(objects.size() * 2);
    nodes.push_back(BvhNode{});
    build_recursive(nodes, order, boxes, 0, 0, objects.size());
}

thread_local OpTally* g_tally = nullptr;
bool g_tally_on = false;

// bvh.rs:97-150
void Bvh::query(V3 p, const std::vector<Object>& objects, DistHit& best) const {
    if (empty) return;
    uint32_t stack[64];
    size_t sp = 0;
    stack[sp++] = 0;
    if (g_tally) g_tally->ops += kTallyAabbOps;
    while (sp > 0) {
        sp -= 1;
        size_t ni = stack[sp];
        const BvhNode& n = nodes[ni];
        if (aabb_dist(p, n.lo, n.hi) >= best.d) continue;
        if (n.count > 0) {
            for (uint32_t k = n.start; k < n.start + n.count; ++k) {
                const Object& o = objects[order[k]];
                float d = o.dist(p);
                if (d < best.d) {
                    best.d = d;
                    best.mat = o.mat;
                }
            }
        } else {
            size_t l = ni + 1, r = n.right;
            float dl = aabb_dist(p, nodes[l].lo, nodes[l].hi);
            float dr = aabb_dist(p, nodes[r].lo, nodes[r].hi);
            if (g_tally) g_tally->ops += 2u * kTallyAabbOps;
            if (dl < dr) {
                stack[sp++] = (uint32_t)r;
                stack[sp++] = (uint32_t)l;
            } else {
                stack[sp++] = (uint32_t)l;
                stack[sp++] = (uint32_t)r;
            }
        }
    }
}

// bvh.rs:152-163
void Bvh::query_linear(V3 p, const std::vector<Object>& objects, DistHit& best) const {
    for (const Object& o : objects) {
        float d = o.dist(p);
        if (d < best.d) {
            best.d = d;
            best.mat = o.mat;
        }
    }
}

// EXT: blend groups.  Objects with blend > 0 form one smooth-union group of
// radius k: the group distance is smin_poly(b1, b2, k) of its two nearest
// members.  Pruning uses total + k, which is exact because an object farther
// than b1 + k cannot change the smooth minimum.
static inline void blend_update(const Object& o, float d, DistHit& hard, DistHit& b1, float& b2) {
    if (o.blend > 0.0f) {
        if (d < b1.d) {
            b2 = b1.d;
            b1.d = d;
            b1.mat = o.mat;
        } else if (d < b2) {
            b2 = d;
        }
    } else if (d < hard.d) {
        hard.d = d;
        hard.mat = o.mat;
    }
}

void Bvh::query_blend(V3 p, const std::vector<Object>& objects, float k, DistHit& hard, DistHit& b1,
                      float& b2, bool linear) const {
    if (empty) return;
    if (linear) {
        for (const Object& o : objects) blend_update(o, o.dist(p), hard, b1, b2);
        return;
    }
    uint32_t stack[64];
    size_t sp = 0;
    stack[sp++] = 0;
    if (g_tally) g_tally->ops += kTallyAabbOps;
    while (sp > 0) {
        sp -= 1;
        size_t ni = stack[sp];
        const BvhNode& n = nodes[ni];
        float thr = fmin_(hard.d, smin_poly(b1.d, b2, k)) + k;
        if (aabb_dist(p, n.lo, n.hi) >= thr) continue;
        if (n.count > 0) {
            for (uint32_t q = n.start; q < n.start + n.count; ++q) {
                const Object& o = objects[order[q]];
                blend_update(o, o.dist(p), hard, b1, b2);
            }
        } else {
            size_t l = ni + 1, r = n.right;
            float dl = aabb_dist(p, nodes[l].lo, nodes[l].hi);
            float dr = aabb_dist(p, nodes[r].lo, nodes[r].hi);
            if (g_tally) g_tally->ops += 2u * kTallyAabbOps;
            if (dl < dr) {
                stack[sp++] = (uint32_t)r;
                stack[sp++] = (uint32_t)l;
            } else {
                stack[sp++] = (uint32_t)l;
                stack[sp++] = (uint32_t)r;
            }
        }
    }
}

// ---------------------------------------------------------------- scene.rs:191-236
DistHit Scene::dist_mat(V3 p) const {
    DistHit best{FLT_MAX, 0};
    if (g_tally) g_tally->ops += kTallyEvalOps;
    for (const Object& o : planes) {
        float d = o.dist(p);
        if (d < best.d) best = DistHit{d, o.mat};
    }
    if (blend_k <= 0.0f) {
        if (use_bvh)
            bvh.query(p, objects, best);
        else
            bvh.query_linear(p, objects, best);
        return best;
    }
    DistHit b1{FLT_MAX, 0};
    float b2 = FLT_MAX;
    bvh.query_blend(p, objects, blend_k, best, b1, b2, !use_bvh);
    float s = smin_poly(b1.d, b2, blend_k);
    if (s < best.d) return DistHit{s, b1.mat};
    return best;
}

V3 Scene::normal(V3 p) const {
    const float H = 5e-4f;
    static const V3 K[4] = {{1.0f, -1.0f, -1.0f}, {-1.0f, -1.0f, 1.0f}, {-1.0f, 1.0f, -1.0f}, {1.0f, 1.0f, 1.0f}};
    V3 n = splat(0.0f);
    for (int i = 0; i < 4; ++i) n += K[i] * dist(p + K[i] * H);
    float l = length(n);
    if (l > 1e-12f) return n / l;
    return v3(0.0f, 1.0f, 0.0f);
}

// ---------------------------------------------------------------- scene.rs:94-161
static const uint32_t MAX_DIMENSION = 16384;
static const uint64_t MAX_PIXELS = 40000000ull;
static const uint64_t MAX_SURFELS = 2000000ull;
static const uint32_t MAX_BOUNCES = 4096;

bool RenderCfg::validate(std::string& err) const {
    std::ostringstream os;
    if (width == 0 || height == 0) {
        err = "width and height must be at least 1";
        return false;
    }
    if (width > MAX_DIMENSION || height > MAX_DIMENSION) {
        os << "image is " << width << "x" << height << ", which exceeds the " << MAX_DIMENSION
           << " px limit on either side";
        err = os.str();
        return false;
    }
    uint64_t px = (uint64_t)width * (uint64_t)height;
    if (px > MAX_PIXELS) {
        os << "image is " << px << " pixels (" << width << "x" << height << "); the limit is " << MAX_PIXELS;
        err = os.str();
        return false;
    }
    if (surfels == 0) {
        err = "surfels must be at least 1";
        return false;
    }
    if (surfels > MAX_SURFELS) {
        os << "surfels is " << surfels << "; the limit is " << MAX_SURFELS;
        err = os.str();
        return false;
    }
    if (bounces == 0) {
        err = "bounces must be at least 1";
        return false;
    }
    if (bounces > MAX_BOUNCES) {
        os << "bounces is " << bounces << "; the limit is " << MAX_BOUNCES;
        err = os.str();
        return false;
    }
    if (!std::isfinite(exposure) || exposure <= 0.0f) {
        os << "exposure must be a positive finite number, got " << exposure;
        err = os.str();
        return false;
    }
    return true;
}

// ---------------------------------------------------------------- scene.rs:247-300
static bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

static std::vector<std::string> tokenize(const std::string& src) {
    std::vector<std::string> out;
    std::string cur;
    bool in_str = false;
    size_t i = 0;
    while (i < src.size()) {
        char c = src[i++];
        if (in_str) {
            if (c == '"') {
                out.push_back("\"" + cur);
                cur.clear();
                in_str = false;
            } else {
                cur.push_back(c);
            }
            continue;
        }
        if (c == '#') {
            while (i < src.size()) {
                if (src[i++] == '\n') break;
            }
            continue;
        }
        if (c == '"') {
            if (!cur.empty()) {
                out.push_back(cur);
                cur.clear();
            }
            in_str = true;
            continue;
        }
        if (c == '{' || c == '}' || c == '[' || c == ']' || c == ',' || c == ':') {
            if (!cur.empty()) {
                out.push_back(cur);
                cur.clear();
            }
            out.push_back(std::string(1, c));
            continue;
        }
        if (is_ws(c)) {
            if (!cur.empty()) {
                out.push_back(cur);
                cur.clear();
            }
            continue;
        }
        cur.push_back(c);
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

namespace {
struct ParseError {
    std::string msg;
};

struct P {
    std::vector<std::string> t;
    size_t i = 0;
    const std::string* peek() const { return i < t.size() ? &t[i] : nullptr; }
    std::string next_tok() {
        if (i >= t.size()) throw ParseError{"unexpected end of file"};
        return t[i++];
    }
    void expect(const std::string& s) {
        std::string n = next_tok();
        if (n != s) throw ParseError{"expected `" + s + "`, found `" + n + "`"};
    }
    float fl() {
        std::string n = next_tok();
        // Rust's f32 grammar: optional sign, decimal digits / inf / nan, no hex,
        // no surrounding whitespace; the whole token must be consumed.
        bool ok = !n.empty();
        size_t s = (ok && (n[0] == '+' || n[0] == '-')) ? 1 : 0;
        if (ok && n.size() > s + 1 && n[s] == '0' && (n[s + 1] == 'x' || n[s + 1] == 'X')) ok = false;
        float v = 0.0f;
        if (ok) {
            char* end = nullptr;
            v = std::strtof(n.c_str(), &end);
            ok = end && *end == '\0' && end != n.c_str();
            if (ok && n.find('(') != std::string::npos) ok = false;
        }
        if (!ok) throw ParseError{"expected number, found `" + n + "`"};
        return v;
    }
    void skip_comma() {
        if (peek() && *peek() == ",") i += 1;
    }
    V3 vec3() {
        expect("[");
        float x = fl();
        skip_comma();
        float y = fl();
        skip_comma();
        float z = fl();
        skip_comma();
        expect("]");
        return v3(x, y, z);
    }
    std::string ident() {
        std::string n = next_tok();
        size_t k = 0;
        while (k < n.size() && n[k] == '"') ++k;
        return n.substr(k);
    }
    bool key(std::string& k) {
        if (peek() && *peek() == "}") {
            i += 1;
            return false;
        }
        k = next_tok();
        expect(":");
        return true;
    }
};
}  // namespace

// scene.rs:573-601 (EXT: sphere and point lights also enclose the bounds)
static void world_bounds(const Scene& sc, V3& out_lo, V3& out_hi) {
    V3 lo = splat(FLT_MAX), hi = splat(-FLT_MAX);
    for (const Object& o : sc.objects) {
        V3 a, b;
        if (o.prim.bounds(a, b)) {
            lo = vmin(lo, a);
            hi = vmax(hi, b);
        }
    }
    for (const QuadLight& l : sc.lights) {
        for (int k = 0; k < 4; ++k) {
            lo = vmin(lo, l.verts[k]);
            hi = vmax(hi, l.verts[k]);
        }
    }
    for (const SphereLight& l : sc.sphere_lights) {
        lo = vmin(lo, l.c - splat(l.r));
        hi = vmax(hi, l.c + splat(l.r));
    }
    for (const PointLight& l : sc.point_lights) {
        lo = vmin(lo, l.pos);
        hi = vmax(hi, l.pos);
    }
    if (lo.x > hi.x) {
        out_lo = splat(-5.0f);
        out_hi = splat(5.0f);
        return;
    }
    bool has_planes = !sc.planes.empty();
    float diag = length(hi - lo);
    float pad = has_planes ? diag * 0.25f : diag * 0.02f;
    out_lo = lo - splat(pad);
    out_hi = hi + splat(pad);
}

// scene.rs:365-571
bool parse_scene(const std::string& src, Scene& sc, std::string& err) {
    P p;
    p.t = tokenize(src);
    sc = Scene{};
    sc.materials.push_back(Material{});
    std::map<std::string, uint32_t> mat_ids;
    try {
        while (p.peek()) {
            std::string block = *p.peek();
            p.i += 1;
            std::string k;
            if (block == "camera") {
                p.expect("{");
                while (p.key(k)) {
                    if (k == "pos") sc.camera.pos = p.vec3();
                    else if (k == "look") sc.camera.look = p.vec3();
                    else if (k == "up") sc.camera.up = p.vec3();
                    else if (k == "fov") sc.camera.fov = p.fl();
                    else throw ParseError{"unknown camera key `" + k + "`"};
                    p.skip_comma();
                }
            } else if (block == "render") {
                p.expect("{");
                while (p.key(k)) {
                    if (k == "width") sc.cfg.width = as_u32(p.fl());
                    else if (k == "height") sc.cfg.height = as_u32(p.fl());
                    else if (k == "surfels") sc.cfg.surfels = as_u64(p.fl());
                    else if (k == "bounces") sc.cfg.bounces = as_u32(p.fl());
                    else if (k == "exposure") sc.cfg.exposure = p.fl();
                    else throw ParseError{"unknown render key `" + k + "`"};
                    p.skip_comma();
                }
            } else if (block == "material") {
                std::string name = p.ident();
                Material m;
                p.expect("{");
                while (p.key(k)) {
                    if (k == "albedo") m.albedo = p.vec3();
                    else if (k == "roughness") m.roughness = clampf(p.fl(), 0.02f, 1.0f);
                    else if (k == "metallic") m.metallic = clampf(p.fl(), 0.0f, 1.0f);
                    else if (k == "emissive") m.emissive = p.vec3();
                    else throw ParseError{"unknown material key `" + k + "`"};
                    p.skip_comma();
                }
                mat_ids[name] = (uint32_t)sc.materials.size();
                sc.materials.push_back(m);
            } else if (block == "light") {
                QuadLight l;
                for (int q = 0; q < 4; ++q) l.verts[q] = splat(0.0f);
                l.emit = splat(10.0f);
                p.expect("{");
                while (p.key(k)) {
                    if (k == "verts") {
                        p.expect("[");
                        for (int q = 0; q < 4; ++q) {
                            l.verts[q] = p.vec3();
                            p.skip_comma();
                        }
                        p.expect("]");
                    } else if (k == "emit") {
                        l.emit = p.vec3();
                    } else {
                        throw ParseError{"unknown light key `" + k + "`"};
                    }
                    p.skip_comma();
                }
                sc.lights.push_back(l);
            } else if (block == "sphere_light") {  // EXT
                SphereLight l{splat(0.0f), 0.25f, splat(10.0f)};
                p.expect("{");
                while (p.key(k)) {
                    if (k == "center") l.c = p.vec3();
                    else if (k == "radius") l.r = p.fl();
                    else if (k == "emit") l.emit = p.vec3();
                    else throw ParseError{"unknown sphere_light key `" + k + "`"};
                    p.skip_comma();
                }
                if (!(l.r > 0.0f)) throw ParseError{"sphere_light radius must be positive"};
                sc.sphere_lights.push_back(l);
            } else if (block == "point_light") {  // EXT
                PointLight l{splat(0.0f), splat(10.0f)};
                p.expect("{");
                while (p.key(k)) {
                    if (k == "pos") l.pos = p.vec3();
                    else if (k == "emit") l.emit = p.vec3();
                    else throw ParseError{"unknown point_light key `" + k + "`"};
                    p.skip_comma();
                }
                sc.point_lights.push_back(l);
            } else if (block == "sphere" || block == "box" || block == "plane" || block == "capsule" ||
                       block == "torus" || block == "cylinder" || block == "cone") {
                uint32_t mat = 0;
                std::vector<Prim> subtract;
                V3 center = splat(0.0f);
                float radius = 1.0f;
                V3 size = splat(1.0f);
                float round = 0.0f;
                V3 normal = v3(0.0f, 1.0f, 0.0f);
                float height = 0.0f;
                V3 a = splat(0.0f);
                V3 b = v3(0.0f, 1.0f, 0.0f);
                float major = 1.0f, minor = 0.25f, radius_top = 0.0f;
                float blend = 0.0f;
                p.expect("{");
                while (p.key(k)) {
                    if (k == "center") center = p.vec3();
                    else if (k == "radius") radius = p.fl();
                    else if (k == "size") size = p.vec3();
                    else if (k == "round") round = p.fl();
                    else if (k == "normal") normal = normalize(p.vec3());
                    else if (k == "height") height = p.fl();
                    else if (k == "a") a = p.vec3();
                    else if (k == "b") b = p.vec3();
                    else if (k == "major") major = p.fl();
                    else if (k == "minor") minor = p.fl();
                    else if (k == "radius_top") radius_top = p.fl();
                    else if (k == "mat") {
                        std::string n = p.ident();
                        auto it = mat_ids.find(n);
                        if (it == mat_ids.end()) throw ParseError{"unknown material `" + n + "`"};
                        mat = it->second;
                    } else if (k == "subtract_sphere") {
                        p.expect("[");
                        V3 c = p.vec3();
                        p.skip_comma();
                        float r = p.fl();
                        p.skip_comma();
                        p.expect("]");
                        Prim s{};
                        s.kind = SPHERE;
                        s.c = c;
                        s.r = r;
                        subtract.push_back(s);
                    } else if (k == "subtract_cylinder") {  // EXT
                        p.expect("[");
                        V3 ca = p.vec3();
                        p.skip_comma();
                        V3 cb = p.vec3();
                        p.skip_comma();
                        float r = p.fl();
                        p.skip_comma();
                        p.expect("]");
                        Prim s{};
                        s.kind = CYLINDER;
                        s.a = ca;
                        s.b = cb;
                        s.r = r;
                        subtract.push_back(s);
                    } else if (k == "blend") {  // EXT
                        blend = fmax_(p.fl(), 0.0f);
                    } else {
                        throw ParseError{"unknown shape key `" + k + "`"};
                    }
                    p.skip_comma();
                }
                Prim pr{};
                if (block == "sphere") {
                    pr.kind = SPHERE;
                    pr.c = center;
                    pr.r = radius;
                } else if (block == "box") {
                    pr.kind = ROUNDBOX;
                    pr.c = center;
                    pr.b = size * 0.5f;
                    pr.r = fmin_(round, min_elem(size) * 0.499f);
                } else if (block == "plane") {
                    pr.kind = PLANE;
                    pr.n = normal;
                    pr.h = height;
                } else if (block == "capsule") {
                    pr.kind = CAPSULE;
                    pr.a = a;
                    pr.b = b;
                    pr.r = radius;
                } else if (block == "cylinder") {
                    pr.kind = CYLINDER;
                    pr.a = a;
                    pr.b = b;
                    pr.r = radius;
                } else if (block == "cone") {
                    pr.kind = CONE;
                    pr.a = a;
                    pr.b = b;
                    pr.ra = radius;
                    pr.rb = radius_top;
                } else {
                    pr.kind = TORUS;
                    pr.c = center;
                    pr.major = major;
                    pr.minor = minor;
                }
                Object obj;
                obj.prim = pr;
                obj.subtract = subtract;
                obj.mat = mat;
                obj.blend = (pr.kind == PLANE) ? 0.0f : blend;
                if (pr.kind == PLANE)
                    sc.planes.push_back(obj);
                else
                    sc.objects.push_back(obj);
            } else {
                throw ParseError{"unknown block `" + block + "`"};
            }
        }
    } catch (const ParseError& e) {
        err = e.msg;
        return false;
    }
    if (sc.lights.empty() && sc.sphere_lights.empty() && sc.point_lights.empty()) {
        err = "scene has no `light` block";
        return false;
    }
    if (!sc.cfg.validate(err)) return false;
    sc.blend_k = 0.0f;
    for (const Object& o : sc.objects) sc.blend_k = fmax_(sc.blend_k, o.blend);
    sc.bvh.build(sc.objects);
    world_bounds(sc, sc.lo, sc.hi);
    return true;
}

}  // namespace orc
