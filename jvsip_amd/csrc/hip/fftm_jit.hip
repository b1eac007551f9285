// fftm_jit.hip -- kernels compiled at plan creation for one schedule.
//
// A plan whose passes the generic kernel would run (complex FP32 or FP64;
// rows, columns or the passes of a split) gets the same Stockham schedule compiled with every
// constant fixed (fftm_static.hpp): hipRTC compiles the template instance
// named by the schedule for the device's architecture, the code object is
// loaded as a module, and the function is kept in a process-wide cache keyed
// by the schedule, so later plans of the same schedule and every execution
// find it without compiling.  The device headers travel inside the library
// (gen_jit_sources.py).  A schedule that does not compile or load is
// remembered as unavailable and runs on the generic kernel.
//
// With VSIP_AMD_JIT_CACHE_DIR set, compiled code objects are also kept on
// disk, one file per schedule, and later processes load them instead of
// compiling (the file name hashes the instance, the architecture, hipRTC's
// version and the embedded device sources, so a changed library or device
// never loads a stale object).
//
// VSIP_AMD_JIT=0 turns the path off.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <unistd.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>
#include "vhip.h"

extern "C" {
extern const int vhip_jit_nheaders;
extern const char *const vhip_jit_header_names[];
extern const char *const vhip_jit_header_texts[];
}

namespace {

struct Entry {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;   // nullptr: the schedule is not available
};

std::mutex g_mu;
std::map<std::string, Entry> g_cache;
char g_err[512];

// The launch path looks kernels up by the schedule's fields packed into
// integers (no instance-name formatting per launch); prepare fills it.
struct FastKey {
    unsigned long long a, b;
    bool operator<(const FastKey &o) const { return a != o.a ? a < o.a : b < o.b; }
};
std::map<FastKey, hipFunction_t> g_fast;
// launch-time compiles already attempted (each key once; a failure stays on
// the generic kernel without another attempt)
std::map<FastKey, bool> g_tried;
// whether a launch may compile a layout its plan did not prepare
bool g_launch_compiles = true;

// every field in its own bits: n (31), vpb (16), threads (11), prec (4) in a;
// min (6), mout (5), nst (4) and radix - 1 of each stage (4 each, radices
// 1..16) in b; a schedule the fields cannot hold is not keyed (it runs on
// the generic kernel)
bool fast_key(int prec, int n, int nst, const int *radix, int vpb, int threads, int min, int mout, FastKey &k) {
    if (n < 1 || vpb < 1 || vpb > 0xffff || threads < 1 || threads > 2047 || nst < 1 || nst > 8) return false;
    k.a = (unsigned long long)(unsigned)n | ((unsigned long long)vpb << 31) | ((unsigned long long)threads << 47) |
          ((unsigned long long)(prec & 0xf) << 58);
    if (min < 0 || min > 0x3f || mout < 0 || mout > 0x1f) return false;
    k.b = (unsigned long long)min | ((unsigned long long)mout << 6) | ((unsigned long long)nst << 11);
    for (int s = 0; s < nst; ++s) {
        if (radix[s] < 1 || radix[s] > 16) return false;
        k.b |= (unsigned long long)(radix[s] - 1) << (15 + 4 * s);
    }
    return true;
}

bool jit_on() {
    const char *e = getenv("VSIP_AMD_JIT");
    return !(e && e[0] == '0');
}

// the template instance of a schedule; empty when the schedule has no
// compiled form (more than 8 stages, or not a single pass)
std::string instance(int prec, int n, int nst, const int *radix, int vpb, int threads, int min, int mout) {
    // load-side bits: ACROSS, FUSED, REAL, C2R, SPLIT_IN; store-side: ACROSS,
    // FUSED, PTW, SPLIT_OUT
    const int min_bits = VHIP_JIT_ACROSS | VHIP_JIT_FUSED | VHIP_JIT_REAL | VHIP_JIT_C2R | VHIP_JIT_SPLIT_IN;
    if (nst < 1 || nst > 8 || min < 0 || (min & ~min_bits) || mout < 0 || mout > 31 || (prec != 4 && prec != 8))
        return "";
    // split sides: plain complex data only (no taper, no real side, no
    // accumulation, no post-twiddle)
    if ((min & VHIP_JIT_SPLIT_IN) && (min & (VHIP_JIT_FUSED | VHIP_JIT_REAL | VHIP_JIT_C2R))) return "";
    if ((mout & VHIP_JIT_SPLIT_OUT) && ((mout & (VHIP_JIT_FUSED | VHIP_JIT_PTW)) || (min & (VHIP_JIT_REAL | VHIP_JIT_C2R))))
        return "";
    if ((mout & VHIP_JIT_PTW) && (mout & VHIP_JIT_FUSED)) return "";  // no power accumulation in a split pass
    if ((min & VHIP_JIT_REAL) && (mout & VHIP_JIT_FUSED)) return "";   // no power accumulation after a real input
    // complex to real: nothing fused, no post-twiddle, not also real input
    if ((min & VHIP_JIT_C2R) && ((min & (VHIP_JIT_FUSED | VHIP_JIT_REAL)) || mout > 1)) return "";
    int r[8];
    for (int s = 0; s < 8; ++s) r[s] = s < nst ? radix[s] : 1;
    char buf[256];
    snprintf(buf, sizeof buf, "vhipk::st::fftm_static_kernel<%s,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d>",
             prec == 4 ? "float" : "double", n, vpb, threads, min, mout, nst, r[0], r[1], r[2], r[3], r[4], r[5],
             r[6], r[7]);
    return buf;
}

// layout mode of one side of a descriptor: 0 contiguous rows (single-level
// vectors), 1 across vectors (any strides, span known); -1 neither
int side_mode(const vhip_view &g, int map, long nvec) {
    // split sides are addressed per vector offset and element stride (any
    // strides within the span); the mode is then the lane mapping only
    if (g.split) return g.span > 0 ? (map == VHIP_MAP_ACROSS ? 1 : 0) : -1;
    if (map == VHIP_MAP_ALONG && g.contig && g.nv0 >= nvec) return 0;
    if (map == VHIP_MAP_ACROSS && g.span > 0) return 1;
    return -1;
}

std::string arch_option() {
    std::string a = "gfx950";
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) {
        a = p.gcnArchName;
        size_t c = a.find(':');
        if (c != std::string::npos) a = a.substr(0, c);
    }
    return "--offload-arch=" + a;
}

// compile `name`; code object into `code`, lowered name into `lowered`
int compile(const std::string &name, const std::string &arch, std::vector<char> &code, std::string &lowered) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, "#include \"fftm_static.hpp\"\n", "fftm_jit.hip", vhip_jit_nheaders,
                            vhip_jit_header_texts, vhip_jit_header_names) != HIPRTC_SUCCESS)
        return -1;
    hiprtcAddNameExpression(prog, name.c_str());
    const char *opts[] = {arch.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops"};
    hiprtcResult r = hiprtcCompileProgram(prog, 7, opts);
    int rc = 0;
    if (r != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::vector<char> log(ls + 1, 0);
        hiprtcGetProgramLog(prog, log.data());
        snprintf(g_err, sizeof g_err, "%s: %.400s", hiprtcGetErrorString(r), log.data());
        rc = -2;
    } else {
        const char *low = nullptr;
        size_t cs = 0;
        if (hiprtcGetLoweredName(prog, name.c_str(), &low) != HIPRTC_SUCCESS || !low ||
            hiprtcGetCodeSize(prog, &cs) != HIPRTC_SUCCESS || cs == 0) {
            rc = -3;
        } else {
            lowered = low;
            code.resize(cs);
            if (hiprtcGetCode(prog, code.data()) != HIPRTC_SUCCESS) rc = -3;
        }
    }
    hiprtcDestroyProgram(&prog);
    return rc;
}

// ---- code objects on disk (VSIP_AMD_JIT_CACHE_DIR) ----------------------

uint64_t fnv1a(uint64_t h, const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
    return h;
}

uint64_t sources_hash() {
    static const uint64_t h = [] {
        uint64_t v = 0xcbf29ce484222325ull;
        for (int i = 0; i < vhip_jit_nheaders; ++i) {
            v = fnv1a(v, vhip_jit_header_names[i], strlen(vhip_jit_header_names[i]) + 1);
            v = fnv1a(v, vhip_jit_header_texts[i], strlen(vhip_jit_header_texts[i]) + 1);
        }
        return v;
    }();
    return h;
}

// the cache file of one instance; empty when the cache is off
std::string cache_path(const std::string &name, const std::string &arch) {
    const char *dir = getenv("VSIP_AMD_JIT_CACHE_DIR");
    if (!dir || !*dir) return std::string();
    int maj = 0, mnr = 0;
    hiprtcVersion(&maj, &mnr);
    uint64_t h = sources_hash();
    h = fnv1a(h, name.c_str(), name.size() + 1);
    h = fnv1a(h, arch.c_str(), arch.size() + 1);
    h = fnv1a(h, &maj, sizeof maj);
    h = fnv1a(h, &mnr, sizeof mnr);
    char file[48];
    snprintf(file, sizeof file, "/vsip_amd_%016llx.co", (unsigned long long)h);
    return std::string(dir) + file;
}

// file layout: a line "vsip_amd jit 1 <lowered name>", then the code object
const char kMagic[] = "vsip_amd jit 1 ";

bool cache_read(const std::string &path, std::vector<char> &code, std::string &lowered) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    std::vector<char> all;
    char buf[65536];
    size_t got;
    while ((got = fread(buf, 1, sizeof buf, f)) > 0) all.insert(all.end(), buf, buf + got);
    fclose(f);
    const size_t m = sizeof kMagic - 1;
    if (all.size() <= m || memcmp(all.data(), kMagic, m) != 0) return false;
    size_t nl = m;
    while (nl < all.size() && all[nl] != '\n') ++nl;
    if (nl + 1 >= all.size()) return false;
    lowered.assign(all.data() + m, all.data() + nl);
    code.assign(all.begin() + (long)nl + 1, all.end());
    // only an ELF image goes to the loader; anything else is compiled anew
    return !lowered.empty() && code.size() > 64 && memcmp(code.data(), "\x7f" "ELF", 4) == 0;
}

// written under a temporary name and renamed, so that no process reads a
// partial file; a failure leaves no file and is not an error
void cache_write(const std::string &path, const std::vector<char> &code, const std::string &lowered) {
    const std::string tmp = path + ".tmp." + std::to_string((long)getpid());
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return;
    bool ok = fwrite(kMagic, 1, sizeof kMagic - 1, f) == sizeof kMagic - 1 &&
              fwrite(lowered.data(), 1, lowered.size(), f) == lowered.size() && fputc('\n', f) != EOF &&
              fwrite(code.data(), 1, code.size(), f) == code.size();
    ok = (fclose(f) == 0) && ok;
    if (!ok || rename(tmp.c_str(), path.c_str()) != 0) remove(tmp.c_str());
}

// load a code object and find the instance's function in it
bool load_module(const std::vector<char> &code, const std::string &lowered, Entry &e) {
    if (hipModuleLoadData(&e.mod, code.data()) != hipSuccess) {
        e.mod = nullptr;
        return false;
    }
    if (hipModuleGetFunction(&e.fn, e.mod, lowered.c_str()) != hipSuccess) {
        (void)hipModuleUnload(e.mod);
        e.mod = nullptr;
        e.fn = nullptr;
        return false;
    }
    return true;
}

}  // namespace

// Compile (no device needed) the kernel of one schedule for `arch`
// ("gfx950"); 0 on success.  Used to check the embedded sources.
extern "C" int vhip_jit_compile_check(int prec, int n, int nst, const int *radix, int vpb, int threads, int min,
                                      int mout, const char *arch, size_t *code_bytes) {
    std::string name = instance(prec, n, nst, radix, vpb, threads, min, mout);
    if (name.empty()) return -1;
    std::vector<char> code;
    std::string lowered;
    int rc = compile(name, std::string("--offload-arch=") + arch, code, lowered);
    if (code_bytes) *code_bytes = code.size();
    return rc;
}

extern "C" const char *vhip_jit_last_error(void) { return g_err; }

// Make the kernel of a schedule available (compile and load once); 0 when
// it is, nonzero when the schedule runs on the generic kernel.
extern "C" int vhip_jit_prepare(int prec, int n, int nst, const int *radix, int vpb, int threads, int min,
                                int mout) {
    if (!jit_on()) return -1;
    std::string name = instance(prec, n, nst, radix, vpb, threads, min, mout);
    if (name.empty()) return -1;
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_cache.find(name);
    if (it != g_cache.end()) return it->second.fn ? 0 : -1;
    Entry e;
    std::vector<char> code;
    std::string lowered;
    const std::string arch = arch_option();
    const std::string path = cache_path(name, arch);
    if (path.empty() || !cache_read(path, code, lowered) || !load_module(code, lowered, e)) {
        if (compile(name, arch, code, lowered) == 0 && load_module(code, lowered, e) && !path.empty())
            cache_write(path, code, lowered);
    }
    (void)hipGetLastError();
    g_cache[name] = e;
    FastKey key;
    if (e.fn && fast_key(prec, n, nst, radix, vpb, threads, min, mout, key)) g_fast[key] = e.fn;
    return e.fn ? 0 : -1;
}

// Is the kernel of this schedule compiled and loaded?
extern "C" int vhip_jit_ready(int prec, int n, int nst, const int *radix, int vpb, int threads, int min, int mout) {
    std::string name = instance(prec, n, nst, radix, vpb, threads, min, mout);
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_cache.find(name);
    return (it != g_cache.end() && it->second.fn) ? 1 : 0;
}

// layouts of a descriptor as template arguments (bit 1: fused taper / power
// accumulation); false when the compiled kernels cannot run it
static bool desc_layouts(const vhip_fftm_desc *d, int &min, int &mout) {
    if ((d->prec != 4 && d->prec != 8) || (d->kind != VHIP_C2C && d->kind != VHIP_R2C && d->kind != VHIP_C2R))
        return false;
    if (d->kind == VHIP_R2C && (d->acc || d->ptw)) return false;
    if (d->kind == VHIP_C2R && (d->acc || d->ptw || d->tap)) return false;
    if ((d->in.split || d->out.split) && d->kind != VHIP_C2C) return false;
    if ((d->in.split && d->tap) || (d->out.split && (d->acc || d->ptw))) return false;
    min = side_mode(d->in, d->map_in, d->nvec);
    mout = side_mode(d->out, d->map_out, d->nvec);
    if (min < 0 || mout < 0) return false;
    if (d->tap) min |= VHIP_JIT_FUSED;
    if (d->acc) mout |= VHIP_JIT_FUSED;
    if (d->ptw) mout |= VHIP_JIT_PTW;
    if (d->kind == VHIP_R2C) min |= VHIP_JIT_REAL;
    if (d->kind == VHIP_C2R) min |= VHIP_JIT_C2R;
    if (d->in.split) min |= VHIP_JIT_SPLIT_IN;
    if (d->out.split) mout |= VHIP_JIT_SPLIT_OUT;
    return true;
}

extern "C" void vhip_jit_launch_compiles(int on) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_launch_compiles = on != 0;
}

// Compile the kernel a descriptor needs (for launches the plan did not
// prepare at creation, such as fused pipelines); 0 when it is available.
extern "C" int vhip_jit_prepare_desc(const vhip_fftm_desc *d) {
    int min, mout;
    if (!desc_layouts(d, min, mout)) return -1;
    return vhip_jit_prepare(d->prec, d->n, d->nst, d->radix, d->vpb, d->threads, min, mout);
}

// Launch d on its compiled kernel when one is loaded and d has the layout
// the kernel serves; -100 otherwise (the caller uses the generic kernel).
// A plan compiles the layouts its major implies at creation; a view with
// another one (a COL plan on contiguous columns, a ROW plan across
// column-major storage) has its kernel compiled at its first launch, once.
extern "C" int vhip_jit_try_launch(const vhip_fftm_desc *d, void *stream) {
    int min, mout;
    if (!d->compiled || !desc_layouts(d, min, mout)) return -100;
    const bool real = (min & VHIP_JIT_REAL) != 0, c2r = (min & VHIP_JIT_C2R) != 0;
    const long es = 2L * d->prec;
    // contiguous sides: rows must not overlap and a group's rows span < 2 GiB
    // (real sides in real elements)
    const long in_len = real ? 2L * d->n : c2r ? d->n + 1L : d->n, in_esz = real ? d->prec : es;
    const long out_len = real ? d->n + 1L : c2r ? 2L * d->n : d->n, out_esz = c2r ? d->prec : es;
    if ((min & 1) == 0) {
        if (d->nvec > 1 && d->in.vs0 < in_len) return -100;
        if (((long)(d->vpb - 1) * d->in.vs0 + in_len) * in_esz >= (1L << 31)) return -100;
    }
    if ((mout & 1) == 0) {
        if (d->nvec > 1 && d->out.vs0 < out_len) return -100;
        if (((long)(d->vpb - 1) * d->out.vs0 + out_len) * out_esz >= (1L << 31)) return -100;
    }
    hipFunction_t fn = nullptr;
    FastKey key;
    if (!fast_key(d->prec, d->n, d->nst, d->radix, d->vpb, d->threads, min, mout, key)) return -100;
    bool compile_now = false;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        auto it = g_fast.find(key);
        if (it != g_fast.end()) fn = it->second;
        else if (!g_launch_compiles || g_tried.find(key) != g_tried.end()) return -100;
        else compile_now = g_tried[key] = true;
    }
    if (compile_now) {
        // the kernel's static LDS (the group's padded vectors and offsets) must fit
        int lds_max = 65536;
        vhip_lds_max(&lds_max);
        if ((long)d->vpb * d->lds_stride * 2 * d->prec + (long)d->vpb * 12 > lds_max) return -100;
        if (vhip_jit_prepare(d->prec, d->n, d->nst, d->radix, d->vpb, d->threads, min, mout) != 0) return -100;
        std::lock_guard<std::mutex> lock(g_mu);
        auto it = g_fast.find(key);
        if (it == g_fast.end()) return -100;
        fn = it->second;
    }
    const long grid = (d->nvec + d->vpb - 1) / d->vpb;
    if (grid > 0x7fffffffL) return -100;
    vhip_fftm_desc arg = *d;
    void *args[] = {&arg};
    hipError_t e = hipModuleLaunchKernel(fn, (unsigned)grid, 1, 1, (unsigned)d->threads, 1, 1, 0,
                                         (hipStream_t)stream, args, nullptr);
    return (int)e;
}
