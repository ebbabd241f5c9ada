// The physim transform-plugin protocol over the GPU engine: `astro`,
// `astro2` and `simple_astro` as C-ABI transform elements. The symbol set and
// semantics follow the reference's register_plugin! macro
// (physim-core/src/plugin/mod.rs:118-160), the #[transform_element] proc
// macro (physim-attribute/src/lib.rs:39-202) and the C plugin example
// (c_plugin/plugin.c:25-190).
#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include <set>

#include "pg_engine.h"
#include "physim_abi.h"
#include "physim_gpu.h"

namespace {

// ---- a strict JSON reader, enough for an element's property object --------
// Accepts exactly one JSON value of any kind; the caller keeps top-level
// numeric members. Invalid JSON fails, as serde_json::from_str does in the
// reference's init (physim-attribute/src/lib.rs:67-70).
struct Json {
    const char* p;
    const char* end;

    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool lit(const char* s) {
        size_t n = strlen(s);
        if ((size_t)(end - p) < n || memcmp(p, s, n) != 0) return false;
        p += n;
        return true;
    }
    bool string(std::string* out) {
        if (p >= end || *p != '"') return false;
        ++p;
        while (p < end && *p != '"') {
            char c = *p++;
            if ((unsigned char)c < 0x20) return false;
            if (c == '\\') {
                if (p >= end) return false;
                char esc = *p++;
                switch (esc) {
                    case '"': case '\\': case '/': if (out) *out += esc; break;
                    case 'b': if (out) *out += '\b'; break;
                    case 'f': if (out) *out += '\f'; break;
                    case 'n': if (out) *out += '\n'; break;
                    case 'r': if (out) *out += '\r'; break;
                    case 't': if (out) *out += '\t'; break;
                    case 'u':
                        for (int k = 0; k < 4; ++k) {
                            if (p >= end || !isxdigit((unsigned char)*p)) return false;
                            ++p;
                        }
                        if (out) *out += '?';
                        break;
                    default: return false;
                }
            } else if (out) {
                *out += c;
            }
        }
        if (p >= end) return false;
        ++p;
        return true;
    }
    bool number(double* out) {
        const char* s = p;
        if (p < end && *p == '-') ++p;
        if (p >= end) return false;
        if (*p == '0') {
            ++p;
        } else if (*p >= '1' && *p <= '9') {
            while (p < end && isdigit((unsigned char)*p)) ++p;
        } else {
            return false;
        }
        if (p < end && *p == '.') {
            ++p;
            if (p >= end || !isdigit((unsigned char)*p)) return false;
            while (p < end && isdigit((unsigned char)*p)) ++p;
        }
        if (p < end && (*p == 'e' || *p == 'E')) {
            ++p;
            if (p < end && (*p == '+' || *p == '-')) ++p;
            if (p >= end || !isdigit((unsigned char)*p)) return false;
            while (p < end && isdigit((unsigned char)*p)) ++p;
        }
        if (out) {
            std::string tmp(s, p);
            *out = strtod(tmp.c_str(), nullptr);
        }
        return true;
    }
    // kind: 0 other, 1 number, 2 string
    bool value(int depth, int* kind, double* num, std::string* str) {
        if (depth > 128) return false;
        ws();
        if (p >= end) return false;
        if (kind) *kind = 0;
        char c = *p;
        if (c == '{') {
            ++p;
            ws();
            if (p < end && *p == '}') { ++p; return true; }
            for (;;) {
                ws();
                if (!string(nullptr)) return false;
                ws();
                if (p >= end || *p != ':') return false;
                ++p;
                if (!value(depth + 1, nullptr, nullptr, nullptr)) return false;
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == '}') { ++p; return true; }
                return false;
            }
        }
        if (c == '[') {
            ++p;
            ws();
            if (p < end && *p == ']') { ++p; return true; }
            for (;;) {
                if (!value(depth + 1, nullptr, nullptr, nullptr)) return false;
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == ']') { ++p; return true; }
                return false;
            }
        }
        if (c == '"') {
            if (kind) *kind = 2;
            return string(str);
        }
        if (c == 't') return lit("true");
        if (c == 'f') return lit("false");
        if (c == 'n') return lit("null");
        if (kind) *kind = 1;
        return number(num);
    }
};

struct Props {
    bool has_theta = false, has_e = false;
    double theta = 0.0, e = 0.0;
    std::string precision;
};

// Parses a JSON object of properties; false on invalid JSON or a non-object.
bool parse_props(const uint8_t* buf, size_t len, Props* out) {
    Json j{(const char*)buf, (const char*)buf + len};
    j.ws();
    if (j.p >= j.end || *j.p != '{') return false;
    ++j.p;
    j.ws();
    if (j.p < j.end && *j.p == '}') {
        ++j.p;
    } else {
        for (;;) {
            j.ws();
            std::string key;
            if (!j.string(&key)) return false;
            j.ws();
            if (j.p >= j.end || *j.p != ':') return false;
            ++j.p;
            int kind = 0;
            double num = 0.0;
            std::string str;
            if (!j.value(1, &kind, &num, &str)) return false;
            if (key == "theta") {  // last duplicate wins, as in a HashMap insert
                out->has_theta = kind == 1;
                out->theta = num;
            } else if (key == "e") {
                out->has_e = kind == 1;
                out->e = num;
            } else if (key == "precision" && kind == 2) {
                out->precision = str;
            }
            j.ws();
            if (j.p < j.end && *j.p == ',') { ++j.p; continue; }
            if (j.p < j.end && *j.p == '}') { ++j.p; break; }
            return false;
        }
    }
    j.ws();
    return j.p == j.end;
}

// ---- element instances ------------------------------------------------------
struct Element {
    pg_params params;
    const char* name;
};

std::mutex g_mu;
pg_engine* g_engine = nullptr;
void* g_bus = nullptr;
std::set<const void*> g_elements;  // live gravity elements (for the verlet binding)

pg_engine* engine() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_engine) {
        int dev = 0;
        if (const char* s = getenv("PHYSIM_GPU_DEVICE")) dev = atoi(s);
        if (pg_engine_create(dev, 0, &g_engine) != PG_OK) {
            fprintf(stderr, "physim_gpu: cannot create the GPU engine: %s. Aborting\n",
                    pg_last_error());
            abort();
        }
    }
    return g_engine;
}

// transformers.rs:162-180 (astro2), :71-89 (astro), :246-256 (simple_astro):
// theta defaults to 1.0; e defaults to 1.0 and is taken by absolute value.
void* element_init(int kind, const char* name, const uint8_t* cfg, uintptr_t len) {
    if (!cfg) return nullptr;
    Props pr;
    if (!parse_props(cfg, len, &pr)) return nullptr;
    Element* el = new Element();
    el->name = name;
    el->params.kind = kind;
    el->params.theta = pr.has_theta ? pr.theta : 1.0;
    el->params.e = pr.has_e ? std::fabs(pr.e) : 1.0;
    el->params.dt = 0.0;
    el->params.precision = pr.precision == "fp64" ? PG_PREC_FP64 : PG_PREC_MIXED;
    std::lock_guard<std::mutex> lk(g_mu);
    g_elements.insert(el);
    return el;
}

// physim-attribute/src/lib.rs:89-98: the reference aborts when its transform
// panics; a GPU failure aborts the same way.
void element_transform(const void* obj, const Entity* state, uintptr_t n, Acceleration* acc,
                       uintptr_t n_acc) {
    if (!obj) return;
    const Element* el = (const Element*)obj;
    if (n == 0) return;
    if (n_acc < n) {
        fprintf(stderr, "Problem encountered in the %s element's transform method "
                        "(acceleration buffer %zu < state %zu). Aborting\n",
                el->name, (size_t)n_acc, (size_t)n);
        abort();
    }
    int rc = pg_transform_aos(engine(), &el->params, state, n, acc);
    if (rc != PG_OK) {
        fprintf(stderr, "Problem encountered in the %s element's transform method: %s. Aborting\n",
                el->name, pg_last_error());
        abort();
    }
}

void element_destroy(void* obj) {
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_elements.erase(obj);
    }
    delete (Element*)obj;
}

char* describe(void* obj, RustStringAllocFn alloc, bool with_theta) {
    if (!obj || !alloc) return nullptr;
    std::string s = "{";
    if (with_theta)
        s += "\"theta\":\"Barnes-Hut parameter. Increase for speed, decrease for accuracy. "
             "Default=1.0\",";
    s += "\"e\":\"Easing factor. Modify G*Ma*Mb*(r-e)^-2. Default=1.0\",";
    s += "\"precision\":\"mixed (FP32 pair forces, FP64 tree and sums of record) or fp64. "
         "Default=mixed\"}";
    return alloc(s.c_str());
}

void element_recv(void* obj, const CMessage* msg) {
    (void)obj;
    (void)msg;
}

typedef void (*PostBusFn)(void*, CMessage);

// transformers.rs:198-203: every gravity element announces itself to the
// energy sink. post_bus_callback is exported by the host (physim.h:81); a
// host without it only loses the announcements. A GPU element also tells a
// device-resident integrator (the verlet shim, which receives bus messages
// as every element does: physim-core/src/messages.rs) that its transform
// can run on the device: topic "physim_gpu", message "transform", the
// element as sender_id, which the shim hands to pg_verlet_bind.
void element_post_config(void* obj) {
    if (!obj || !g_bus) return;
    PostBusFn post = (PostBusFn)dlsym(RTLD_DEFAULT, "post_bus_callback");
    if (!post) return;
    CMessage m;
    m.priority = PHYSIM_PRIO_LOW;
    m.topic = "energysink";
    m.message = "gravity";
    m.sender_id = (uintptr_t)obj;
    m.origin = 1;
    post(g_bus, m);
    m.topic = "physim_gpu";
    m.message = "transform";
    post(g_bus, m);
}

ElementMetaFFI meta(RustStringAllocFn alloc, const char* name, const char* blurb) {
    ElementMetaFFI m;
    m.kind = PHYSIM_TRANSFORM;
    m.name = alloc(name);
    m.plugin = alloc("physim_gpu");
    m.version = alloc("0.1.0");
    m.license = alloc("MIT");
    m.author = alloc("physim_gpu authors");
    m.blurb = alloc(blurb);
    m.repo = alloc("https://github.com/jhb123/physim");
    return m;
}

#define PG_ELEMENT(NAME, KIND, WITH_THETA, BLURB)                                          \
    void* NAME##_init_impl(const uint8_t* c, uintptr_t l) {                                \
        return element_init(KIND, #NAME, c, l);                                            \
    }                                                                                      \
    char* NAME##_props_impl(void* o, RustStringAllocFn a) { return describe(o, a, WITH_THETA); } \
    const TransformElementAPI NAME##_api = {NAME##_init_impl, element_transform,           \
                                            element_destroy,  NAME##_props_impl,           \
                                            element_recv,     element_post_config};

PG_ELEMENT(astro2, PG_KIND_OCTREE, true, "")
PG_ELEMENT(astro, PG_KIND_QUADTREE, true, "")
PG_ELEMENT(simple_astro, PG_KIND_DIRECT, false, "")

}  // namespace

pg_engine* pg_plugin_engine(void) { return engine(); }

bool pg_plugin_element_params(const void* element, pg_params* out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!element || !g_elements.count(element)) return false;
    *out = ((const Element*)element)->params;
    return true;
}

extern "C" {

const char* get_plugin_abi_info(void) { return "C"; }

const char* register_plugin(void) { return "astro,astro2,simple_astro"; }

void set_callback_target(void* target) {
    if (target == NULL) {
        fprintf(stderr, "Error: callback target is null\n");
        abort();
    }
    g_bus = target;
}

const TransformElementAPI* astro2_get_api(void) { return &astro2_api; }
const TransformElementAPI* astro_get_api(void) { return &astro_api; }
const TransformElementAPI* simple_astro_get_api(void) { return &simple_astro_api; }

ElementMetaFFI astro2_register(RustStringAllocFn alloc) {
    return meta(alloc, "astro2",
                "Compute approximate gravitational accelerations with the Barnes-Hut algorithm "
                "(octree) on the GPU");
}
ElementMetaFFI astro_register(RustStringAllocFn alloc) {
    return meta(alloc, "astro",
                "Compute approximate gravitational accelerations with the Barnes-Hut algorithm "
                "(quadtree) on the GPU");
}
ElementMetaFFI simple_astro_register(RustStringAllocFn alloc) {
    return meta(alloc, "simple_astro", "Compute exact gravitational accelerations on the GPU");
}

}  // extern "C"
