// config.hip -- the reference's JSON network configuration, read into
// insta_config and insta_run_config (host code only).
//
// Testbed::load_network_config + merge_parent_network_config
// (instant-ngp/src/testbed.cu:77-152): a JSON file (comments allowed) whose
// "parent" key names another file, relative to the child's directory, that is
// loaded first (recursively) and then merge-patched with the child (RFC 7386:
// objects merge key by key, null removes a key, anything else replaces).
// The merged document is then read the way Testbed::reset_network
// (testbed.cu:2181-2300), NerfNetwork's constructor (nerf_network.h:135-152)
// and rta::Core::load_training_data (src/core.cu:199-221) read it.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "context.hpp"

namespace insta {
namespace {

struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;  // insertion order

    const Json* get(const char* key) const {
        if (kind != Object) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    Json* find(const std::string& key) {
        for (auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

// recursive-descent parser; // and /* */ comments are whitespace
class Parser {
public:
    explicit Parser(const std::string& s) : s_(s) {}
    bool parse(Json& out, std::string& err) {
        if (!value(out, 0) || (skip(), i_ != s_.size())) {
            err = err_.empty() ? "trailing characters" : err_;
            err += " at byte " + std::to_string(i_);
            return false;
        }
        return true;
    }

private:
    const std::string& s_;
    size_t i_ = 0;
    std::string err_;

    bool fail(const char* m) {
        if (err_.empty()) err_ = m;
        return false;
    }
    void skip() {
        for (;;) {
            while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '/') {
                while (i_ < s_.size() && s_[i_] != '\n') ++i_;
            } else if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '*') {
                const size_t e = s_.find("*/", i_ + 2);
                i_ = e == std::string::npos ? s_.size() : e + 2;
            } else {
                return;
            }
        }
    }
    bool literal(const char* w) {
        const size_t n = std::strlen(w);
        if (s_.compare(i_, n, w) != 0) return fail("bad literal");
        i_ += n;
        return true;
    }
    bool string(std::string& out) {
        if (s_[i_] != '"') return fail("expected a string");
        ++i_;
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\') {
                if (i_ >= s_.size()) return fail("bad escape");
                const char e = s_[i_++];
                switch (e) {
                    case 'n': c = '\n'; break;
                    case 't': c = '\t'; break;
                    case 'r': c = '\r'; break;
                    case 'b': c = '\b'; break;
                    case 'f': c = '\f'; break;
                    case 'u': {  // code points below 0x80 only (keys and values here are ASCII)
                        if (i_ + 4 > s_.size()) return fail("bad \\u escape");
                        const unsigned long cp = std::strtoul(s_.substr(i_, 4).c_str(), nullptr, 16);
                        i_ += 4;
                        if (cp >= 0x80) return fail("non-ASCII \\u escape");
                        c = (char)cp;
                        break;
                    }
                    default: c = e;
                }
            }
            out.push_back(c);
        }
        if (i_ >= s_.size()) return fail("unterminated string");
        ++i_;
        return true;
    }
    bool value(Json& v, int depth) {
        if (depth > 64) return fail("nesting too deep");
        skip();
        if (i_ >= s_.size()) return fail("unexpected end");
        const char c = s_[i_];
        if (c == '{') {
            v.kind = Json::Object;
            ++i_;
            skip();
            if (i_ < s_.size() && s_[i_] == '}') {
                ++i_;
                return true;
            }
            for (;;) {
                skip();
                std::string key;
                if (!string(key)) return false;
                skip();
                if (i_ >= s_.size() || s_[i_] != ':') return fail("expected ':'");
                ++i_;
                Json item;
                if (!value(item, depth + 1)) return false;
                if (Json* prev = v.find(key))
                    *prev = std::move(item);  // a repeated key keeps the last value
                else
                    v.obj.emplace_back(std::move(key), std::move(item));
                skip();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == '}') {
                    ++i_;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = Json::Array;
            ++i_;
            skip();
            if (i_ < s_.size() && s_[i_] == ']') {
                ++i_;
                return true;
            }
            for (;;) {
                Json item;
                if (!value(item, depth + 1)) return false;
                v.arr.push_back(std::move(item));
                skip();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == ']') {
                    ++i_;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Json::String;
            return string(v.str);
        }
        if (c == 't') {
            v.kind = Json::Bool;
            v.b = true;
            return literal("true");
        }
        if (c == 'f') {
            v.kind = Json::Bool;
            v.b = false;
            return literal("false");
        }
        if (c == 'n') {
            v.kind = Json::Null;
            return literal("null");
        }
        const char* start = s_.c_str() + i_;
        char* end = nullptr;
        errno = 0;
        const double d = std::strtod(start, &end);
        if (end == start || errno == ERANGE) return fail("bad number");
        v.kind = Json::Number;
        v.num = d;
        i_ += (size_t)(end - start);
        return true;
    }
};

// RFC 7386 merge patch (nlohmann::json::merge_patch)
void merge_patch(Json& target, const Json& patch) {
    if (patch.kind != Json::Object) {
        target = patch;
        return;
    }
    if (target.kind != Json::Object) target = Json{Json::Object};
    for (const auto& kv : patch.obj) {
        if (kv.second.kind == Json::Null) {
            for (auto it = target.obj.begin(); it != target.obj.end(); ++it)
                if (it->first == kv.first) {
                    target.obj.erase(it);
                    break;
                }
        } else if (Json* t = target.find(kv.first)) {
            merge_patch(*t, kv.second);
        } else {
            Json fresh;
            merge_patch(fresh, kv.second);
            target.obj.emplace_back(kv.first, std::move(fresh));
        }
    }
}

bool read_file(const std::string& path, std::string& out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) out.append(buf, n);
    std::fclose(f);
    return true;
}

std::string parent_dir(const std::string& path) {
    const size_t k = path.find_last_of('/');
    return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

// merge_parent_network_config (testbed.cu:77-89)
bool load_merged(const std::string& path, Json& out, std::string& err, int depth) {
    if (depth > 16) {
        err = "config: \"parent\" chain deeper than 16 files at " + path;
        return false;
    }
    std::string text;
    if (!read_file(path, text)) {
        err = "config: cannot read " + path;
        return false;
    }
    Json doc;
    std::string perr;
    if (!Parser(text).parse(doc, perr)) {
        err = "config: " + path + ": " + perr;
        return false;
    }
    const Json* parent = doc.get("parent");
    if (!parent) {
        out = std::move(doc);
        return true;
    }
    if (parent->kind != Json::String) {
        err = "config: \"parent\" is not a string in " + path;
        return false;
    }
    Json base;
    if (!load_merged(parent_dir(path) + parent->str, base, err, depth + 1)) return false;
    merge_patch(base, doc);
    out = std::move(base);
    return true;
}

bool lower_eq(const std::string& a, const char* b) {
    if (a.size() != std::strlen(b)) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (std::tolower((unsigned char)a[i]) != std::tolower((unsigned char)b[i])) return false;
    return true;
}

bool num(const Json* j, double& out) {
    if (!j) return false;
    if (j->kind == Json::Number) {
        out = j->num;
        return true;
    }
    if (j->kind == Json::Bool) {
        out = j->b ? 1.0 : 0.0;
        return true;
    }
    return false;
}
template <class T>
void take(const Json* parent, const char* key, T& field) {
    double v;
    if (parent && num(parent->get(key), v)) field = (T)v;
}

}  // namespace

insta_status load_config_json(const char* path, insta_config* cfg, insta_run_config* run) {
    Json doc;
    std::string err;
    if (!load_merged(path, doc, err, 0)) {
        set_error(err);
        return INSTA_ERR_INVALID;
    }
    if (doc.kind != Json::Object) {
        set_error("config: the document is not a JSON object");
        return INSTA_ERR_INVALID;
    }
    insta_config c;
    insta_default_config(&c);
    insta_run_config r;
    insta_default_run_config(&r);
    auto unsupported = [&](const std::string& what) {
        set_error("config: " + what + " is not supported by this library");
        return INSTA_ERR_UNSUPPORTED;
    };

    // "encoding": the hash grid (reset_network, testbed.cu:2224-2268)
    if (const Json* e = doc.get("encoding")) {
        const Json* ot = e->get("otype");
        if (ot && ot->kind == Json::String && !lower_eq(ot->str, "HashGrid"))
            return unsupported("encoding otype \"" + ot->str + "\"");
        uint32_t F = 2;
        take(e, "n_features_per_level", F);
        c.n_features_per_level = F;
        double nf = 0.0;
        if (num(e->get("n_features"), nf) && nf > 0.0)
            c.n_levels = (uint32_t)nf / F;
        else {
            c.n_levels = 16;
            take(e, "n_levels", c.n_levels);
        }
        c.log2_hashmap_size = 15;
        take(e, "log2_hashmap_size", c.log2_hashmap_size);
        c.base_resolution = 0;
        take(e, "base_resolution", c.base_resolution);
        if (!c.base_resolution) c.base_resolution = 1u << (c.log2_hashmap_size / 3);
        float pls = 0.0f;
        take(e, "per_level_scale", pls);
        if (pls <= 0.0f && c.n_levels > 1)  // finest level 2048 over the unit cube, aabb_scale 1
            pls = std::exp(std::log(2048.0f * 1.0f / (float)c.base_resolution) / (float)(c.n_levels - 1));
        c.per_level_scale = pls;
    }
    // "network" (density MLP) and "rgb_network" (nerf_network.h:140-152)
    auto mlp = [&](const Json* n, uint32_t& hidden, uint32_t* out_dims) -> insta_status {
        if (!n) return INSTA_OK;
        const Json* ot = n->get("otype");
        if (ot && ot->kind == Json::String && !lower_eq(ot->str, "FullyFusedMLP") && !lower_eq(ot->str, "CutlassMLP") &&
            !lower_eq(ot->str, "MegakernelMLP"))
            return unsupported("network otype \"" + ot->str + "\"");
        const Json* act = n->get("activation");
        if (act && act->kind == Json::String && !lower_eq(act->str, "ReLU"))
            return unsupported("activation \"" + act->str + "\"");
        const Json* oact = n->get("output_activation");
        if (oact && oact->kind == Json::String && !lower_eq(oact->str, "None"))
            return unsupported("output_activation \"" + oact->str + "\"");
        uint32_t width = c.n_neurons;
        take(n, "n_neurons", width);
        if (width != c.n_neurons) return unsupported("n_neurons other than " + std::to_string(c.n_neurons));
        take(n, "n_hidden_layers", hidden);
        if (out_dims) take(n, "n_output_dims", *out_dims);
        return INSTA_OK;
    };
    if (const Json* n = doc.get("network")) {
        uint32_t w = 64;
        take(n, "n_neurons", w);
        c.n_neurons = w;
    }
    insta_status st = mlp(doc.get("network"), c.density_hidden_layers, &c.density_out);
    if (st != INSTA_OK) return st;
    st = mlp(doc.get("rgb_network"), c.rgb_hidden_layers, nullptr);
    if (st != INSTA_OK) return st;
    // "dir_encoding": spherical harmonics, alone or first in a Composite
    if (const Json* d = doc.get("dir_encoding")) {
        const Json* sh = d;
        const Json* ot = d->get("otype");
        if (ot && ot->kind == Json::String && lower_eq(ot->str, "Composite")) {
            const Json* nested = d->get("nested");
            sh = nested && nested->kind == Json::Array && !nested->arr.empty() ? &nested->arr[0] : nullptr;
            ot = sh ? sh->get("otype") : nullptr;
        }
        if (!ot || ot->kind != Json::String || !lower_eq(ot->str, "SphericalHarmonics"))
            return unsupported("direction encoding other than SphericalHarmonics");
        take(sh, "degree", c.sh_degree);
    }
    // loss: the training step implements Huber (testbed_nerf.cu:1299)
    if (const Json* l = doc.get("loss")) {
        const Json* ot = l->get("otype");
        if (ot && ot->kind == Json::String && !lower_eq(ot->str, "Huber"))
            return unsupported("loss \"" + ot->str + "\"");
    }
    // optimizer: Ema over Adam, or Adam (configs/nerf/main.json:21-32)
    if (const Json* o = doc.get("optimizer")) {
        const Json* adam = o;
        const Json* ot = o->get("otype");
        r.ema_decay = 0.0f;
        if (ot && ot->kind == Json::String && lower_eq(ot->str, "Ema")) {
            r.ema_decay = 0.99f;
            take(o, "decay", r.ema_decay);
            adam = o->get("nested");
            ot = adam ? adam->get("otype") : nullptr;
        }
        if (!adam || !ot || ot->kind != Json::String || !lower_eq(ot->str, "Adam"))
            return unsupported("optimizer other than Adam (optionally under Ema)");
        take(adam, "learning_rate", r.learning_rate);
        take(adam, "beta1", r.beta1);
        take(adam, "beta2", r.beta2);
        take(adam, "epsilon", r.epsilon);
        take(adam, "l2_reg", r.l2_reg);
    }
    // INSTA's run keys (src/core.cu:199-221, configs/nerf/main.json)
    take(&doc, "disable_interp", c.disable_interp);
    take(&doc, "exp_params", r.exp_params);
    take(&doc, "geo_params", r.geo_params);
    take(&doc, "use_eyes", r.use_eyes);
    take(&doc, "optimize_latent_code", r.optimize_latent_code);
    take(&doc, "max_steps", r.max_training_steps);
    take(&doc, "depth_weight", r.depth_weight);
    take(&doc, "seg_weight", r.seg_weight);
    take(&doc, "max_cached_bvh", r.max_cached_bvh);
    take(&doc, "max_images_gpu", r.max_images_gpu);
    // the density input's extra dims are the expression parameters
    // (N_EXP_PARAMS becomes the dataset's n_extra_dims)
    c.n_extra_dims = r.exp_params;
    *cfg = c;
    if (run) *run = r;
    return INSTA_OK;
}

}  // namespace insta

extern "C" {

void insta_default_run_config(insta_run_config* r) {
    if (!r) return;
    *r = insta_run_config{};
    r->learning_rate = 1e-2f;
    r->beta1 = 0.9f;
    r->beta2 = 0.99f;
    r->epsilon = 1e-15f;
    r->l2_reg = 0.0f;
    r->ema_decay = 0.0f;
    r->exp_params = 0;
    r->max_training_steps = 0;
    r->depth_weight = 0.0f;
    r->seg_weight = 0.0f;
}

insta_status insta_load_config_json(const char* path, insta_config* cfg, insta_run_config* run) {
    if (!path || !cfg) {
        insta::set_error("load_config_json: null argument");
        return INSTA_ERR_INVALID;
    }
    return insta::load_config_json(path, cfg, run);
}

}  // extern "C"
