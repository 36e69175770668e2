// tuning.cpp -- the tuning table: per workload, the coherent beamformer's
// workgroup shape and the coherent DMs' FFT lengths, measured on the device
// by sw_pipeline_autotune and kept in a JSON file that sw_pipeline_create
// reads (skyweaver_amd/tuning/gfx950.json beside the library by default).
//
// File format (the subset of JSON this reader accepts: objects, arrays,
// numbers, strings, true/false/null):
//   {"device": "...", "entries": [
//     {"antennas": A, "beams": B, "samples": T, "tscrunch": ts, "stokes": "I",
//      "cb_waves": 4,                     -- 2, 4 or 8 waves of 32 time rows
//      "dms": [...], "fft_len": [...], "share_forward": 0|1,   -- optional
//      ...}]}                              -- other keys are ignored
// An entry applies to a pipeline of the same antennas, beams, gulp length,
// tscrunch and Stokes mode; its FFT lengths apply when its DM list equals
// the pipeline's.  No reference counterpart: the reference's launch shapes
// and its FFT length (the gulp) are fixed.
#include "sw_common.hpp"

#include <dlfcn.h>

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace sw {
namespace {

struct Json {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    double num = 0;
    bool b = false;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;
    const Json* get(const char* k) const
    {
        if(kind != Obj) return nullptr;
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
};

struct Parser {
    const char* s;
    const char* e;
    std::string err;
    void ws()
    {
        while(s < e && std::isspace((unsigned char)*s)) ++s;
    }
    bool fail(const char* m)
    {
        if(err.empty()) err = m;
        return false;
    }
    bool lit(const char* w)
    {
        const size_t n = std::strlen(w);
        if((size_t)(e - s) < n || std::strncmp(s, w, n) != 0) return fail("bad literal");
        s += n;
        return true;
    }
    bool string(std::string& out)
    {
        if(s >= e || *s != '"') return fail("expected a string");
        ++s;
        while(s < e && *s != '"') {
            if(*s == '\\') {
                if(++s >= e) return fail("bad escape");
                const char c = *s;
                out += c == 'n' ? '\n' : c == 't' ? '\t' : c;  // \" \\ \/ and the two controls
            } else {
                out += *s;
            }
            ++s;
        }
        if(s >= e) return fail("unterminated string");
        ++s;
        return true;
    }
    bool value(Json& v, int depth)
    {
        if(depth > 32) return fail("nesting too deep");
        ws();
        if(s >= e) return fail("unexpected end");
        const char c = *s;
        if(c == '{') {
            v.kind = Json::Obj;
            ++s;
            ws();
            if(s < e && *s == '}') {
                ++s;
                return true;
            }
            for(;;) {
                ws();
                std::string k;
                if(!string(k)) return false;
                ws();
                if(s >= e || *s != ':') return fail("expected ':'");
                ++s;
                if(!value(v.obj[k], depth + 1)) return false;
                ws();
                if(s < e && *s == ',') {
                    ++s;
                    continue;
                }
                if(s < e && *s == '}') {
                    ++s;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if(c == '[') {
            v.kind = Json::Arr;
            ++s;
            ws();
            if(s < e && *s == ']') {
                ++s;
                return true;
            }
            for(;;) {
                v.arr.emplace_back();
                if(!value(v.arr.back(), depth + 1)) return false;
                ws();
                if(s < e && *s == ',') {
                    ++s;
                    continue;
                }
                if(s < e && *s == ']') {
                    ++s;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if(c == '"') {
            v.kind = Json::Str;
            return string(v.str);
        }
        if(c == 't') {
            v.kind = Json::Bool;
            v.b = true;
            return lit("true");
        }
        if(c == 'f') {
            v.kind = Json::Bool;
            return lit("false");
        }
        if(c == 'n') return lit("null");
        char* end = nullptr;
        const std::string tok(s, (size_t)std::min<ptrdiff_t>(e - s, 64));
        v.num = std::strtod(tok.c_str(), &end);
        if(end == tok.c_str()) return fail("bad value");
        v.kind = Json::Num;
        s += end - tok.c_str();
        return true;
    }
};

const char* stokes_name(int s)
{
    static const char* n[] = {"I", "Q", "U", "V", "IQUV"};
    return (s >= 0 && s <= 4) ? n[s] : "?";
}

int num_of(const Json* v, int dflt)
{
    return (v && v->kind == Json::Num) ? (int)std::lround(v->num) : dflt;
}

} // namespace

// The table file to read when a pipeline names none: $SKYWEAVER_AMD_TUNING,
// else tuning/gfx950.json in the package directory of this library
// (<package>/_build/libskyweaver_amd.so).
std::string default_tuning_path()
{
    if(const char* env = std::getenv("SKYWEAVER_AMD_TUNING")) return env;
    Dl_info info{};
    if(dladdr(reinterpret_cast<void*>(&default_tuning_path), &info) && info.dli_fname) {
        std::string lib = info.dli_fname;
        const size_t slash = lib.rfind('/');
        const std::string dir = slash == std::string::npos ? "." : lib.substr(0, slash);
        return dir + "/../tuning/gfx950.json";
    }
    return "tuning/gfx950.json";
}

// Look up the entry for a workload; returns SW_OK with found = 0 when the
// file is absent or has no entry for it.
int tuning_lookup(const std::string& path, int A, int B, int T, int ts, int stokes, int ndms, const float* dms,
                  int* found, int* cb_waves, int* fft_len, int* share_forward)
{
    *found = 0;
    std::ifstream f(path);
    if(!f) return SW_OK;
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    Parser ps{text.data(), text.data() + text.size(), {}};
    Json root;
    if(!ps.value(root, 0)) {
        set_error("tuning table " + path + ": " + ps.err);
        return SW_ERR_INVALID;
    }
    const Json* ents = root.get("entries");
    if(!ents || ents->kind != Json::Arr) {
        set_error("tuning table " + path + ": no \"entries\" array");
        return SW_ERR_INVALID;
    }
    // an entry of this workload whose DM list is the pipeline's, else the
    // first entry of this workload
    const Json* first = nullptr;
    for(const Json& en : ents->arr) {
        const Json* st = en.get("stokes");
        if(num_of(en.get("antennas"), -1) != A || num_of(en.get("beams"), -1) != B ||
           num_of(en.get("samples"), -1) != T || num_of(en.get("tscrunch"), -1) != ts || !st ||
           st->kind != Json::Str || st->str != stokes_name(stokes))
            continue;
        if(!first) first = &en;
        const Json* edms = en.get("dms");
        const Json* efl = en.get("fft_len");
        bool same = edms && efl && edms->kind == Json::Arr && efl->kind == Json::Arr && dms &&
                    (int)edms->arr.size() == ndms && (int)efl->arr.size() == ndms;
        for(int d = 0; same && d < ndms; ++d)
            same = edms->arr[d].kind == Json::Num && std::fabs(edms->arr[d].num - (double)dms[d]) < 1e-6;
        if(same) {
            *found = 2;
            if(cb_waves) *cb_waves = num_of(en.get("cb_waves"), 0);
            if(fft_len)
                for(int d = 0; d < ndms; ++d) fft_len[d] = num_of(&efl->arr[d], 0);
            if(share_forward) *share_forward = num_of(en.get("share_forward"), 0);
            return SW_OK;
        }
    }
    if(first) {
        *found = 1;
        if(cb_waves) *cb_waves = num_of(first->get("cb_waves"), 0);
    }
    return SW_OK;
}

} // namespace sw

extern "C" int sw_tuning_lookup(const char* path, int nantennas, int nbeams, int nsamples, int tscrunch,
                                int stokes, int ndms, const float* dms, int* found, int* cb_waves, int* fft_len,
                                int* share_forward)
{
    SW_CHECK_ARG(found, "null argument");
    SW_CHECK_ARG(ndms >= 0 && (ndms == 0 || dms), "bad DM list");
    const std::string p = path ? std::string(path) : sw::default_tuning_path();
    return sw::tuning_lookup(p, nantennas, nbeams, nsamples, tscrunch, stokes, ndms, dms, found, cb_waves, fft_len,
                             share_forward);
}
