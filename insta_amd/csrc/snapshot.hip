// snapshot.hip -- save/load a trained avatar as a msgpack snapshot.
//
// Restates Testbed::save_snapshot / load_snapshot (instant-ngp/src/
// testbed.cu:2916-3012) for this path: the network configuration as JSON-like
// maps ("encoding", "network", "rgb_network", "dir_encoding", plus "insta":
// every insta_config field) and a "snapshot" map with the reference's keys --
// version, params_type + params_binary (here the FP32 parameters in the
// insta_set_params order), density_grid_size + density_grid_binary (FP16,
// Morton order, empty when the grid was never updated), nerf.rgb ray-batch
// counters, training_step, aabb, render_aabb -- and, when asked, the Adam
// moments.  tiny-cuda-nn's params_binary layout is not in the reference tree,
// so parameters are in this library's order, not the reference's.
// Loading into a context needs the same network shape; the density grid's
// mean and bitfield are rebuilt into a caller-named occupancy slot, as the
// reference's load rebuilds them (update_density_grid_mean_and_bitfield).
//
// The msgpack coder is a small bounds-checked subset: nil, bool, integers,
// float32/64, str, bin, array, map (no ext types).
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "context.hpp"

namespace insta {
namespace {

constexpr uint32_t kSnapshotVersion = 1;
constexpr int kMaxDepth = 32;

// ---- msgpack values ---------------------------------------------------------
struct MpValue {
    enum Kind { Nil, Bool, Int, Float, Str, Bin, Arr, Map } kind = Nil;
    bool b = false;
    int64_t i = 0;
    uint64_t u = 0;  // integers: i and u both set when representable
    bool is_unsigned = false;
    double f = 0.0;
    std::string s;   // Str and Bin payload
    std::vector<MpValue> a;
    std::vector<std::pair<std::string, MpValue>> m;

    const MpValue* get(const char* key) const {
        if (kind != Map) return nullptr;
        for (const auto& kv : m)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    bool number(double& out) const {
        if (kind == Float) out = f;
        else if (kind == Int) out = is_unsigned ? (double)u : (double)i;
        else return false;
        return true;
    }
    bool uint(uint64_t& out) const {
        if (kind != Int || (!is_unsigned && i < 0)) return false;
        out = is_unsigned ? u : (uint64_t)i;
        return true;
    }
};

// ---- writer ------------------------------------------------------------------
class MpWriter {
public:
    std::vector<uint8_t> out;

    void map(uint32_t n) { header(n, 0x80, 16, 0xde, 0xdf); }
    void arr(uint32_t n) { header(n, 0x90, 16, 0xdc, 0xdd); }
    void str(const std::string& v) {
        const size_t n = v.size();
        if (n < 32) byte(0xa0 | (uint8_t)n);
        else if (n < 256) { byte(0xd9); byte((uint8_t)n); }
        else if (n < 65536) { byte(0xda); be(n, 2); }
        else { byte(0xdb); be(n, 4); }
        out.insert(out.end(), v.begin(), v.end());
    }
    void bin(const void* p, size_t n) {
        if (n < 256) { byte(0xc4); byte((uint8_t)n); }
        else if (n < 65536) { byte(0xc5); be(n, 2); }
        else { byte(0xc6); be(n, 4); }
        const uint8_t* q = static_cast<const uint8_t*>(p);
        out.insert(out.end(), q, q + n);
    }
    void uint(uint64_t v) {
        if (v < 128) byte((uint8_t)v);
        else if (v < 256) { byte(0xcc); byte((uint8_t)v); }
        else if (v < 65536) { byte(0xcd); be(v, 2); }
        else if (v <= 0xffffffffull) { byte(0xce); be(v, 4); }
        else { byte(0xcf); be(v, 8); }
    }
    void f32(float v) {
        uint32_t bits;
        std::memcpy(&bits, &v, 4);
        byte(0xca);
        be(bits, 4);
    }
    void boolean(bool v) { byte(v ? 0xc3 : 0xc2); }
    void nil() { byte(0xc0); }
    void key(const char* k) { str(k); }
    void vec3(const float* v) {
        arr(3);
        for (int i = 0; i < 3; ++i) f32(v[i]);
    }

private:
    void byte(uint8_t b) { out.push_back(b); }
    void be(uint64_t v, int n) {
        for (int k = n - 1; k >= 0; --k) out.push_back((uint8_t)(v >> (8 * k)));
    }
    void header(uint32_t n, uint8_t fix, uint32_t fix_max, uint8_t c16, uint8_t c32) {
        if (n < fix_max) byte(fix | (uint8_t)n);
        else if (n < 65536) { byte(c16); be(n, 2); }
        else { byte(c32); be(n, 4); }
    }
};

// ---- reader ------------------------------------------------------------------
class MpReader {
public:
    MpReader(const uint8_t* p, size_t n) : p_(p), n_(n) {}

    bool parse(MpValue& v, int depth = 0) {
        if (depth > kMaxDepth) return false;
        uint8_t t;
        if (!byte(t)) return false;
        if (t <= 0x7f) return set_uint(v, t);
        if (t >= 0xe0) return set_int(v, (int8_t)t);
        if ((t & 0xf0) == 0x80) return parse_map(v, t & 0x0f, depth);
        if ((t & 0xf0) == 0x90) return parse_arr(v, t & 0x0f, depth);
        if ((t & 0xe0) == 0xa0) return payload(v, MpValue::Str, t & 0x1f);
        uint64_t n;
        switch (t) {
            case 0xc0: v.kind = MpValue::Nil; return true;
            case 0xc2: v.kind = MpValue::Bool; v.b = false; return true;
            case 0xc3: v.kind = MpValue::Bool; v.b = true; return true;
            case 0xc4: return be(n, 1) && payload(v, MpValue::Bin, n);
            case 0xc5: return be(n, 2) && payload(v, MpValue::Bin, n);
            case 0xc6: return be(n, 4) && payload(v, MpValue::Bin, n);
            case 0xca: {
                if (!be(n, 4)) return false;
                uint32_t bits = (uint32_t)n;
                float f;
                std::memcpy(&f, &bits, 4);
                v.kind = MpValue::Float;
                v.f = f;
                return true;
            }
            case 0xcb: {
                if (!be(n, 8)) return false;
                double d;
                std::memcpy(&d, &n, 8);
                v.kind = MpValue::Float;
                v.f = d;
                return true;
            }
            case 0xcc: return be(n, 1) && set_uint(v, n);
            case 0xcd: return be(n, 2) && set_uint(v, n);
            case 0xce: return be(n, 4) && set_uint(v, n);
            case 0xcf: return be(n, 8) && set_uint(v, n);
            case 0xd0: return be(n, 1) && set_int(v, (int8_t)n);
            case 0xd1: return be(n, 2) && set_int(v, (int16_t)n);
            case 0xd2: return be(n, 4) && set_int(v, (int32_t)n);
            case 0xd3: return be(n, 8) && set_int(v, (int64_t)n);
            case 0xd9: return be(n, 1) && payload(v, MpValue::Str, n);
            case 0xda: return be(n, 2) && payload(v, MpValue::Str, n);
            case 0xdb: return be(n, 4) && payload(v, MpValue::Str, n);
            case 0xdc: return be(n, 2) && parse_arr(v, n, depth);
            case 0xdd: return be(n, 4) && parse_arr(v, n, depth);
            case 0xde: return be(n, 2) && parse_map(v, n, depth);
            case 0xdf: return be(n, 4) && parse_map(v, n, depth);
            default: return false;  // ext types and reserved codes
        }
    }
    bool at_end() const { return pos_ == n_; }

private:
    const uint8_t* p_;
    size_t n_, pos_ = 0;

    bool byte(uint8_t& b) {
        if (pos_ >= n_) return false;
        b = p_[pos_++];
        return true;
    }
    bool be(uint64_t& v, int n) {
        if (n_ - pos_ < (size_t)n) return false;
        v = 0;
        for (int k = 0; k < n; ++k) v = (v << 8) | p_[pos_++];
        return true;
    }
    bool payload(MpValue& v, MpValue::Kind kind, uint64_t n) {
        if (n > n_ - pos_) return false;
        v.kind = kind;
        v.s.assign(reinterpret_cast<const char*>(p_ + pos_), (size_t)n);
        pos_ += (size_t)n;
        return true;
    }
    static bool set_uint(MpValue& v, uint64_t x) {
        v.kind = MpValue::Int;
        v.u = x;
        v.i = (int64_t)x;
        v.is_unsigned = true;
        return true;
    }
    static bool set_int(MpValue& v, int64_t x) {
        v.kind = MpValue::Int;
        v.i = x;
        v.u = (uint64_t)x;
        v.is_unsigned = x >= 0;
        return true;
    }
    bool parse_arr(MpValue& v, uint64_t n, int depth) {
        if (n > n_ - pos_) return false;  // each element takes at least one byte
        v.kind = MpValue::Arr;
        v.a.resize((size_t)n);
        for (auto& e : v.a)
            if (!parse(e, depth + 1)) return false;
        return true;
    }
    bool parse_map(MpValue& v, uint64_t n, int depth) {
        if (n > (n_ - pos_) / 2) return false;
        v.kind = MpValue::Map;
        v.m.resize((size_t)n);
        for (auto& kv : v.m) {
            MpValue k;
            if (!parse(k, depth + 1) || k.kind != MpValue::Str) return false;
            kv.first = std::move(k.s);
            if (!parse(kv.second, depth + 1)) return false;
        }
        return true;
    }
};

// ---- the configuration as named fields --------------------------------------
struct CfgField {
    const char* name;
    bool is_float;
    size_t offset;
    bool shape;     // must match for parameters to load
    bool optional;  // absent from older snapshots: the default then
};
#define INSTA_U(f, s) {#f, false, offsetof(insta_config, f), s, false}
#define INSTA_F(f, s) {#f, true, offsetof(insta_config, f), s, false}
#define INSTA_UO(f) {#f, false, offsetof(insta_config, f), false, true}
const CfgField kFields[] = {
    INSTA_U(n_levels, true), INSTA_U(n_features_per_level, true), INSTA_U(log2_hashmap_size, true),
    INSTA_U(base_resolution, true), INSTA_F(per_level_scale, true), INSTA_U(n_neurons, true),
    INSTA_U(density_hidden_layers, true), INSTA_U(density_out, true), INSTA_U(rgb_hidden_layers, true),
    INSTA_U(sh_degree, true), INSTA_U(n_extra_dims, true), INSTA_F(density_logit_bias, false),
    INSTA_U(grid_size, true), INSTA_F(dt, false), INSTA_U(max_steps, false), INSTA_F(min_transmittance, false),
    INSTA_U(disable_interp, false), INSTA_U(precision, false), INSTA_U(max_frames, false),
    INSTA_U(max_triangles, false), INSTA_U(max_vertices, false), INSTA_U(bvh_leaf_min, false),
    INSTA_UO(cell_split), INSTA_UO(cell_list_cells), INSTA_UO(cell_list_entries),
};
#undef INSTA_UO
#undef INSTA_U
#undef INSTA_F

void write_config(MpWriter& w, const insta_config& c) {
    w.map((uint32_t)(sizeof(kFields) / sizeof(kFields[0])));
    for (const CfgField& f : kFields) {
        w.key(f.name);
        const char* base = reinterpret_cast<const char*>(&c) + f.offset;
        if (f.is_float) {
            float v;
            std::memcpy(&v, base, 4);
            w.f32(v);
        } else {
            uint32_t v;
            std::memcpy(&v, base, 4);
            w.uint(v);
        }
    }
}

bool read_config(const MpValue& m, insta_config& c) {
    if (m.kind != MpValue::Map) return false;
    insta_config defaults;
    insta_default_config(&defaults);
    for (const CfgField& f : kFields) {
        const MpValue* v = m.get(f.name);
        char* base = reinterpret_cast<char*>(&c) + f.offset;
        if (!v && f.optional) {
            std::memcpy(base, reinterpret_cast<const char*>(&defaults) + f.offset, 4);
            continue;
        }
        if (!v) return false;
        if (f.is_float) {
            double d;
            if (!v->number(d)) return false;
            const float x = (float)d;
            std::memcpy(base, &x, 4);
        } else {
            uint64_t u;
            if (!v->uint(u) || u > 0xffffffffull) return false;
            const uint32_t x = (uint32_t)u;
            std::memcpy(base, &x, 4);
        }
    }
    return true;
}

void write_mlp(MpWriter& w, uint32_t neurons, uint32_t hidden) {
    w.map(5);
    w.key("otype"); w.str("FullyFusedMLP");
    w.key("activation"); w.str("ReLU");
    w.key("output_activation"); w.str("None");
    w.key("n_neurons"); w.uint(neurons);
    w.key("n_hidden_layers"); w.uint(hidden);
}

void write_aabb(MpWriter& w, const float* lo, const float* hi) {
    w.map(2);
    w.key("min"); w.vec3(lo);
    w.key("max"); w.vec3(hi);
}

// NerfDataset as json_binding.h:139-162 writes it (Eigen matrices as arrays of
// rows): the training set's cameras, no file paths (the images are the
// caller's), the render box, identity local frame, unit scale, offset 0.5
void write_dataset(MpWriter& w, const TrainSnapshot& ts, const float* lo, const float* hi) {
    const uint32_t n = ts.n_images;
    w.map(14);
    w.key("n_images"); w.uint(n);
    w.key("paths"); w.arr(n);
    for (uint32_t i = 0; i < n; ++i) w.str("");
    w.key("metadata"); w.arr(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* c = ts.cams.data() + (size_t)i * 16;
        w.map(5);
        w.key("focal_length"); w.arr(2); w.f32(c[12]); w.f32(c[13]);
        w.key("camera_distortion"); w.nil();  // ECameraDistortionMode::None writes nothing
        w.key("principal_point"); w.arr(2); w.f32(c[14]); w.f32(c[15]);
        w.key("rolling_shutter"); w.arr(4); for (int k = 0; k < 4; ++k) w.f32(0.0f);
        w.key("resolution"); w.arr(2); w.uint(ts.width); w.uint(ts.height);
    }
    w.key("xforms"); w.arr(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* c = ts.cams.data() + (size_t)i * 16;
        w.map(2);
        for (const char* k : {"start", "end"}) {
            w.key(k);
            w.arr(3);
            for (int r = 0; r < 3; ++r) {
                w.arr(4);
                for (int q = 0; q < 4; ++q) w.f32(c[4 * r + q]);
            }
        }
    }
    w.key("render_aabb"); write_aabb(w, lo, hi);
    w.key("render_aabb_to_local"); w.arr(3);
    for (int r = 0; r < 3; ++r) {
        w.arr(3);
        for (int q = 0; q < 3; ++q) w.f32(r == q ? 1.0f : 0.0f);
    }
    const float up[3] = {0.0f, 1.0f, 0.0f}, offset[3] = {0.5f, 0.5f, 0.5f};
    w.key("up"); w.vec3(up);
    w.key("offset"); w.vec3(offset);
    w.key("envmap_resolution"); w.arr(2); w.uint(0); w.uint(0);
    w.key("scale"); w.f32(1.0f);
    w.key("aabb_scale"); w.uint(1);
    w.key("from_mitsuba"); w.boolean(false);
    w.key("is_hdr"); w.boolean(false);
    w.key("wants_importance_sampling"); w.boolean(true);
}

bool read_file(const char* path, std::vector<uint8_t>& data) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    bool ok = std::fseek(f, 0, SEEK_END) == 0;
    const long n = ok ? std::ftell(f) : -1;
    ok = ok && n >= 0 && std::fseek(f, 0, SEEK_SET) == 0;
    if (ok) {
        data.resize((size_t)n);
        ok = n == 0 || std::fread(data.data(), 1, (size_t)n, f) == (size_t)n;
    }
    std::fclose(f);
    return ok;
}

insta_status parse_file(const char* path, MpValue& root) {
    std::vector<uint8_t> data;
    if (!read_file(path, data)) {
        set_error(std::string("snapshot: cannot read ") + path);
        return INSTA_ERR_INVALID;
    }
    MpReader r(data.data(), data.size());
    if (!r.parse(root) || !r.at_end() || root.kind != MpValue::Map) {
        set_error("snapshot: not a msgpack map");
        return INSTA_ERR_INVALID;
    }
    return INSTA_OK;
}

}  // namespace
}  // namespace insta

extern "C" {

insta_status insta_save_snapshot(insta_ctx* ctx, const char* path, int include_optimizer_state, void* stream) {
    using namespace insta;
    if (!ctx || !path) {
        set_error("save_snapshot: null argument");
        return INSTA_ERR_INVALID;
    }
    if (!ctx->params_set) {
        set_error("save_snapshot: parameters not set");
        return INSTA_ERR_STATE;
    }
    const hipStream_t s = (hipStream_t)stream;
    const insta_config& c = ctx->cfg;
    const uint64_t np = ctx->layout.n_params;
    std::vector<float> params(np);
    INSTA_HIP(hipMemcpyAsync(params.data(), ctx->d_params, np * sizeof(float), hipMemcpyDeviceToHost, s));
    INSTA_HIP(hipStreamSynchronize(s));
    TrainSnapshot ts;
    insta_status st = train_export(ctx, include_optimizer_state != 0, ts, s);
    if (st != INSTA_OK) return st;
    DensitySnapshot ds;
    st = density_export(ctx, ds, s);
    if (st != INSTA_OK) return st;
    std::vector<_Float16> grid16(ds.grid.size());
    for (size_t i = 0; i < ds.grid.size(); ++i) grid16[i] = (_Float16)ds.grid[i];

    MpWriter w;
    w.map(6);
    w.key("encoding");
    w.map(6);
    w.key("otype"); w.str("HashGrid");
    w.key("n_levels"); w.uint(c.n_levels);
    w.key("n_features_per_level"); w.uint(c.n_features_per_level);
    w.key("log2_hashmap_size"); w.uint(c.log2_hashmap_size);
    w.key("base_resolution"); w.uint(c.base_resolution);
    w.key("per_level_scale"); w.f32(c.per_level_scale);
    w.key("network"); write_mlp(w, c.n_neurons, c.density_hidden_layers);
    w.key("rgb_network"); write_mlp(w, c.n_neurons, c.rgb_hidden_layers);
    w.key("dir_encoding");
    w.map(2);
    w.key("otype"); w.str("SphericalHarmonics");
    w.key("degree"); w.uint(c.sh_degree);
    w.key("insta"); write_config(w, c);
    w.key("snapshot");
    const bool opt = !ts.m.empty();
    w.map(opt ? 13 : 12);
    w.key("version"); w.uint(kSnapshotVersion);
    w.key("n_params"); w.uint(np);
    w.key("params_type"); w.str("float");
    w.key("params_binary"); w.bin(params.data(), np * sizeof(float));
    w.key("density_grid_size"); w.uint(c.grid_size);
    w.key("density_grid_binary"); w.bin(grid16.data(), grid16.size() * sizeof(_Float16));
    w.key("density_grid_state");
    w.map(3);
    w.key("step"); w.uint(ds.step);
    w.key("rng_state"); w.uint(ds.rng_state);
    w.key("rng_inc"); w.uint(ds.rng_inc);
    w.key("nerf");
    w.map(ts.n_images ? 3 : 2);
    w.key("aabb_scale"); w.uint(1);
    w.key("rgb");
    w.map(3);
    w.key("rays_per_batch"); w.uint(ts.present ? ts.rays_per_batch : 1u << 12);
    w.key("measured_batch_size"); w.uint(0);
    w.key("measured_batch_size_before_compaction"); w.uint(0);
    if (ts.n_images) {
        w.key("dataset");
        write_dataset(w, ts, ctx->render_lo, ctx->render_hi);
    }
    w.key("training_step"); w.uint(ts.step);
    w.key("rays_total"); w.uint(ts.rays_total);
    const float unit_lo[3] = {0, 0, 0}, unit_hi[3] = {1, 1, 1};
    w.key("aabb"); write_aabb(w, unit_lo, unit_hi);
    w.key("render_aabb"); write_aabb(w, ctx->render_lo, ctx->render_hi);
    if (opt) {
        w.key("optimizer");
        w.map(4);
        w.key("otype"); w.str("Adam");
        w.key("step"); w.uint(ts.adam_step);
        w.key("first_moments_binary"); w.bin(ts.m.data(), ts.m.size() * sizeof(float));
        w.key("second_moments_binary"); w.bin(ts.v.data(), ts.v.size() * sizeof(float));
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        set_error(std::string("save_snapshot: cannot write ") + path);
        return INSTA_ERR_INVALID;
    }
    const bool ok = std::fwrite(w.out.data(), 1, w.out.size(), f) == w.out.size();
    if (std::fclose(f) != 0 || !ok) {
        set_error(std::string("save_snapshot: write failed for ") + path);
        return INSTA_ERR_INVALID;
    }
    return INSTA_OK;
}

insta_status insta_snapshot_config(const char* path, insta_config* cfg) {
    using namespace insta;
    if (!path || !cfg) {
        set_error("snapshot_config: null argument");
        return INSTA_ERR_INVALID;
    }
    MpValue root;
    insta_status st = parse_file(path, root);
    if (st != INSTA_OK) return st;
    const MpValue* m = root.get("insta");
    insta_config c{};
    if (!m || !read_config(*m, c)) {
        set_error("snapshot_config: no complete \"insta\" configuration map");
        return INSTA_ERR_INVALID;
    }
    *cfg = c;
    return INSTA_OK;
}

insta_status insta_load_snapshot(insta_ctx* ctx, const char* path, int32_t occupancy_slot, void* stream) {
    using namespace insta;
    if (!ctx || !path || occupancy_slot < -1 || occupancy_slot >= (int32_t)ctx->cfg.max_frames) {
        set_error("load_snapshot: null argument or bad occupancy slot");
        return INSTA_ERR_INVALID;
    }
    const hipStream_t s = (hipStream_t)stream;
    MpValue root;
    insta_status st = parse_file(path, root);
    if (st != INSTA_OK) return st;
    const MpValue* snap = root.get("snapshot");
    const MpValue* icfg = root.get("insta");
    insta_config saved{};
    if (!snap || snap->kind != MpValue::Map || !icfg || !read_config(*icfg, saved)) {
        set_error("load_snapshot: file does not contain a snapshot");
        return INSTA_ERR_INVALID;
    }
    uint64_t version = 0;
    const MpValue* ver = snap->get("version");
    if (!ver || !ver->uint(version) || version < kSnapshotVersion) {
        set_error("load_snapshot: snapshot uses an old format");
        return INSTA_ERR_INVALID;
    }
    for (const CfgField& f : kFields) {
        if (!f.shape) continue;
        if (std::memcmp(reinterpret_cast<const char*>(&saved) + f.offset,
                        reinterpret_cast<const char*>(&ctx->cfg) + f.offset, 4) != 0) {
            set_error(std::string("load_snapshot: incompatible network shape (") + f.name + ")");
            return INSTA_ERR_INVALID;
        }
    }
    const uint64_t np = ctx->layout.n_params;
    const MpValue* pt = snap->get("params_type");
    const MpValue* pb = snap->get("params_binary");
    uint64_t n_params = 0;
    const MpValue* pn = snap->get("n_params");
    if (!pt || pt->kind != MpValue::Str || pt->s != "float" || !pb || pb->kind != MpValue::Bin || !pn ||
        !pn->uint(n_params) || n_params != np || pb->s.size() != np * sizeof(float)) {
        set_error("load_snapshot: parameters missing or of another size");
        return INSTA_ERR_INVALID;
    }
    // density grid (FP16 as the reference stores it) and its state
    DensitySnapshot ds;
    const MpValue* gs = snap->get("density_grid_size");
    const MpValue* gb = snap->get("density_grid_binary");
    uint64_t g = 0;
    if (!gs || !gs->uint(g) || g != ctx->cfg.grid_size || !gb || gb->kind != MpValue::Bin) {
        set_error("load_snapshot: incompatible grid size");
        return INSTA_ERR_INVALID;
    }
    const size_t cells = (size_t)g * g * g;
    if (!gb->s.empty()) {
        if (gb->s.size() != cells * sizeof(_Float16)) {
            set_error("load_snapshot: incompatible number of grid cells");
            return INSTA_ERR_INVALID;
        }
        std::vector<_Float16> h(cells);
        std::memcpy(h.data(), gb->s.data(), gb->s.size());
        ds.grid.resize(cells);
        for (size_t i = 0; i < cells; ++i) ds.grid[i] = (float)h[i];
        if (const MpValue* dst = snap->get("density_grid_state")) {
            uint64_t v = 0;
            if (const MpValue* x = dst->get("step"); x && x->uint(v)) ds.step = (uint32_t)v;
            if (const MpValue* x = dst->get("rng_state"); x && x->uint(v)) ds.rng_state = v;
            if (const MpValue* x = dst->get("rng_inc"); x && x->uint(v)) ds.rng_inc = v;
        }
    }
    // training counters and optimizer moments
    TrainSnapshot ts;
    uint64_t v = 0;
    if (const MpValue* x = snap->get("training_step"); x && x->uint(v)) ts.step = (uint32_t)v;
    if (const MpValue* x = snap->get("rays_total"); x && x->uint(v)) ts.rays_total = v;
    ts.rays_per_batch = 1u << 12;
    if (const MpValue* nerf = snap->get("nerf"))
        if (const MpValue* rgb = nerf->get("rgb"))
            if (const MpValue* x = rgb->get("rays_per_batch"); x && x->uint(v) && v > 0 && v <= (1u << 18))
                ts.rays_per_batch = (uint32_t)v;
    if (const MpValue* o = snap->get("optimizer")) {
        const MpValue* m1 = o->get("first_moments_binary");
        const MpValue* m2 = o->get("second_moments_binary");
        if (!m1 || !m2 || m1->kind != MpValue::Bin || m2->kind != MpValue::Bin ||
            m1->s.size() != np * sizeof(float) || m2->s.size() != np * sizeof(float)) {
            set_error("load_snapshot: optimizer state of another size");
            return INSTA_ERR_INVALID;
        }
        ts.m.resize(np);
        ts.v.resize(np);
        if (const MpValue* x = o->get("step"); x && x->uint(v)) ts.adam_step = (uint32_t)v;
        std::memcpy(ts.m.data(), m1->s.data(), np * sizeof(float));
        std::memcpy(ts.v.data(), m2->s.data(), np * sizeof(float));
    }
    // all checks passed: install
    std::vector<float> params(np);
    std::memcpy(params.data(), pb->s.data(), np * sizeof(float));
    INSTA_HIP(hipMemcpyAsync(ctx->d_params, params.data(), np * sizeof(float), hipMemcpyHostToDevice, s));
    INSTA_HIP(hipStreamSynchronize(s));  // pageable source
    reset_inference_average(ctx);
    st = prepare_network(ctx, s);
    if (st != INSTA_OK) return st;
    ctx->params_set = true;
    // counters (and moments, if saved) wait for a later training set
    st = train_import(ctx, ts, s);
    if (st != INSTA_OK) return st;
    return density_import(ctx, ds, (uint32_t)(occupancy_slot + 1), s);
}

}  // extern "C"
