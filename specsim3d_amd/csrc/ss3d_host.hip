// ss3d_host.cpp -- host side of libspecsim3d.so: the C ABI (include/specsim3d.h), grid sizing,
// twiddle tables, spectrum handles, random-stream tables and the pass plans.
//
// Reference counterparts (src/specsim3d.f90): setrot 107-139, cutoff 142-150, nextpow2 181-188,
// sizing 1234-1254, covariance grid + spectrum 1256-1297, seeding 1302-1303, realization loop
// 1305-1344.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/specsim3d.h"
#include "ss3d_exp.h"
#include "ss3d_internal.h"

using namespace ss3d;

// Scratch memory of ss3d_realize_device_noalloc: every work buffer of its passes (fill output,
// y output, realization bases, block statistics) is a window of this array, so that the call
// allocates nothing.  Each GPU holds its own copy from the first use of the library on it until
// the process ends.  In host code `scratch` only names the symbol: its device address comes from
// hipGetSymbolAddress (scratch_address below), taken once per device.
constexpr size_t kScratchBytes = size_t(SS3D_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// ---------------------------------------------------------------- errors
namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(e_ == hipErrorOutOfMemory ? SS3D_ENOMEM : SS3D_EHIP, "%s: %s (%s:%d)", #expr, \
                        hipGetErrorString(e_), __FILE__, __LINE__);                                \
    } while (0)

#define SS3D_TRY(expr)             \
    do {                           \
        int r_ = (expr);           \
        if (r_ != SS3D_OK) return r_; \
    } while (0)

constexpr double kPi = 3.14159265358979323846;  // f90:91
constexpr int64_t kFillBlocks = 2048;  // FP32 strided passes: fewest blocks per launch before members share blocks

// ---------------------------------------------------------------- host model arithmetic
// GSLIB scaled rotation, f90:107-139 (row-major rot, axes).
void setrot(const ss3d_structure& s, double rot[9], double axes[9]) {
    const double deg = kPi / 180.0;
    const double alpha = (s.az >= 0.0 && s.az < 270.0) ? (90.0 - s.az) * deg : (450.0 - s.az) * deg;
    const double beta = -s.dip * deg, theta = s.plg * deg;
    const double sina = std::sin(alpha), cosa = std::cos(alpha);
    const double sinb = std::sin(beta), cosb = std::cos(beta);
    const double sint = std::sin(theta), cost = std::cos(theta);
    const double anis1 = s.ahmin / s.ahmax, anis2 = s.avert / s.ahmax;
    const double afac1 = 1.0 / std::max(anis1, 1.0e-20), afac2 = 1.0 / std::max(anis2, 1.0e-20);
    rot[0] = cosb * cosa;
    rot[1] = cosb * sina;
    rot[2] = -sinb;
    rot[3] = afac1 * (-(cost * sina) + sint * sinb * cosa);
    rot[4] = afac1 * (cost * cosa + sint * sinb * sina);
    rot[5] = afac1 * (sint * cosb);
    rot[6] = afac2 * (sint * sina + cost * sinb * cosa);
    rot[7] = afac2 * (-(sint * cosa) + cost * sinb * sina);
    rot[8] = afac2 * (cost * cosb);
    for (int c = 0; c < 3; ++c) {
        axes[c] = rot[c] * s.ahmax;
        axes[3 + c] = rot[3 + c] / afac1 * s.ahmin;
        axes[6 + c] = rot[6 + c] / afac2 * s.avert;
    }
}

double cutoff(int itype) { return itype == 1 ? 1.0 : (itype == 2 ? 2.3 : 1.6); }  // f90:142-150

int64_t nextpow2(int64_t n) {  // f90:181-188
    int64_t m = 1;
    while (m < n) m *= 2;
    return m;
}

int ilog2(int64_t n) {
    int l = 0;
    while ((int64_t(1) << l) < n) ++l;
    return l;
}

int check_structures(const ss3d_structure* st, int nst) {
    if (!st || nst < 1 || nst > kMaxStruct) return fail(SS3D_EINVAL, "nstruct out of range (1..%d): %d", kMaxStruct, nst);
    for (int s = 0; s < nst; ++s) {
        if (st[s].itype < 1 || st[s].itype > 3)
            return fail(SS3D_EINVAL, "structure %d: itype must be 1, 2 or 3 (got %d)", s + 1, st[s].itype);
        if (!(st[s].ahmax > 0) || !(st[s].ahmin > 0) || !(st[s].avert > 0))
            return fail(SS3D_EINVAL, "ranges must be positive");
    }
    return SS3D_OK;
}

// ---------------------------------------------------------------- Park-Miller on the host
// a b mod (2^31 - 1) for a, b < 2^31 - 1: since 2^31 = 1 (mod M), p = hi 2^31 + lo = hi + lo, one
// conditional subtraction (no division)
uint32_t pm_mul(uint32_t a, uint32_t b) {
    const uint64_t p = (uint64_t)a * b;
    const uint32_t s = (uint32_t)(p & 0x7fffffffu) + (uint32_t)(p >> 31);
    return s >= kPmM ? s - kPmM : s;
}
uint32_t pm_pow(uint32_t a, uint64_t e) {
    e %= (kPmM - 1);  // a^(M-1) = 1, M prime
    uint32_t r = 1, x = a % kPmM;
    while (e) {
        if (e & 1) r = pm_mul(r, x);
        x = pm_mul(x, x);
        e >>= 1;
    }
    return r;
}
// f90:1302-1303: state = mod(|iseed|, 2147483646) + 1, then 10 draws discarded.
uint32_t seed_state(int iseed) {
    int64_t a = iseed < 0 ? -(int64_t)iseed : (int64_t)iseed;
    uint32_t s = (uint32_t)(a % 2147483646LL + 1);
    return pm_mul(s, pm_pow(kPmA, 10));
}

int64_t draws_per_realization(int npx, int npy, int npz, int igauss) {
    int64_t n = (int64_t)npx * npy * npz;
    int64_t nsc = (npx >= 2 ? 2 : 1) * (npy >= 2 ? 2 : 1) * (npz >= 2 ? 2 : 1);
    return (igauss ? 2 : 1) * ((n + nsc) / 2);
}

// ---------------------------------------------------------------- device buffers
// A device allocation owned by the library, or (view) a window of `cap` bytes of memory owned
// elsewhere -- the scratch array of ss3d_realize_device_noalloc -- which never grows or frees.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    bool view = false;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept { swap_with(o); }
    DevBuf& operator=(DevBuf&& o) noexcept {
        swap_with(o);
        return *this;
    }
    ~DevBuf() {
        if (p && !view) (void)hipFree(p);
    }
    void swap_with(DevBuf& o) noexcept {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
        std::swap(view, o.view);
    }
    void set_view(void* base, size_t bytes) {
        p = base;
        cap = bytes;
        view = true;
    }
    int ensure(size_t bytes, hipStream_t s) {
        if (bytes <= cap) return SS3D_OK;
        if (view) return fail(SS3D_ETOOBIG, "scratch window of %zu bytes cannot hold %zu", cap, bytes);
        if (p) {
            HIP_TRY(hipStreamSynchronize(s));
            HIP_TRY(hipFree(p));
            p = nullptr;
            cap = 0;
        }
        HIP_TRY(hipMalloc(&p, bytes));
        cap = bytes;
        return SS3D_OK;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Device buffers handed back by the request loop's evicted spectra and reused by the spectra it
// builds next: a cache miss takes memory from here instead of calling hipMalloc, and an eviction
// calls no hipFree (which waits for the whole device).
struct BufPool {
    std::vector<DevBuf> bufs;
    int64_t bytes = 0;
    // the smallest pooled buffer of at least `need` bytes and at most about twice that (a small
    // table does not take an amplitude array), or -1
    int best(size_t need) const {
        const size_t most = 2 * need + (size_t(1) << 20);
        int k = -1;
        for (int i = 0; i < (int)bufs.size(); ++i)
            if (bufs[i].cap >= need && bufs[i].cap <= most && (k < 0 || bufs[i].cap < bufs[k].cap)) k = i;
        return k;
    }
    // moves a pooled buffer of at least `need` bytes into b (b's own buffer, if any, joins the pool)
    bool take(DevBuf& b, size_t need) {
        const int k = best(need);
        if (k < 0) return false;
        bytes -= (int64_t)bufs[k].cap;
        b.swap_with(bufs[k]);
        if (bufs[k].p) {
            bytes += (int64_t)bufs[k].cap;
        } else {
            bufs[k] = std::move(bufs.back());
            bufs.pop_back();
        }
        return true;
    }
    void put(DevBuf& b) {
        if (!b.p || b.view) return;
        bytes += (int64_t)b.cap;
        bufs.emplace_back();
        bufs.back().swap_with(b);
    }
    // frees the largest pooled buffer (the last resort of the request loop's memory budget)
    bool free_largest() {
        if (bufs.empty()) return false;
        int k = 0;
        for (int i = 1; i < (int)bufs.size(); ++i)
            if (bufs[i].cap > bufs[k].cap) k = i;
        bytes -= (int64_t)bufs[k].cap;
        bufs[k] = std::move(bufs.back());  // the buffer moved out lands in bufs.back() and is freed below
        bufs.pop_back();
        return true;
    }
};

// b holds at least `bytes`: as it is, from the pool, or newly allocated.
int grab(BufPool* pool, DevBuf& b, size_t bytes, hipStream_t s) {
    if (b.cap >= bytes) return SS3D_OK;
    if (pool && pool->take(b, bytes)) return SS3D_OK;
    return b.ensure(bytes, s);
}

// One ordered lane of device work: a stream (plus a second one for the FP32 pipeline) and the
// work buffers a spectrum build or a realization call needs.  A context has one lane; a server
// has one per in-flight request, so their work overlaps on the device.
struct Lane {
    hipStream_t stream = nullptr, stream2 = nullptr;
    DevBuf wa, wb, pipe[2], partial, partial2, rbase, stats, dout, diag;
    hipEvent_t ev_fill[2] = {nullptr, nullptr}, ev_rest[2] = {nullptr, nullptr}, ev_join = nullptr;
    int init() {
        HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
        for (int i = 0; i < 2; ++i) {
            HIP_TRY(hipEventCreateWithFlags(&ev_fill[i], hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&ev_rest[i], hipEventDisableTiming));
        }
        HIP_TRY(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
        return SS3D_OK;
    }
    void sync() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (stream2) (void)hipStreamSynchronize(stream2);
    }
    ~Lane() {
        sync();
        for (auto e : {ev_fill[0], ev_fill[1], ev_rest[0], ev_rest[1], ev_join})
            if (e) (void)hipEventDestroy(e);
        // buffers are freed by their destructors after the streams went idle
    }
    void destroy_streams() {
        sync();
        if (stream) (void)hipStreamDestroy(stream);
        if (stream2) (void)hipStreamDestroy(stream2);
        stream = stream2 = nullptr;
    }
};

__global__ void rbase_kernel(uint32_t s10, uint64_t dmod, uint64_t r0, int n, uint32_t* out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint64_t m1 = kPmM - 1;
    uint64_t r = (r0 + (uint64_t)t) % m1;
    uint64_t e = (r * dmod + 1) % m1;
    uint64_t x = kPmA, acc = 1;
    while (e) {
        if (e & 1) acc = (acc * x) % kPmM;
        x = (x * x) % kPmM;
        e >>= 1;
    }
    out[t] = (uint32_t)((acc * s10) % kPmM);
}

}  // namespace

// ---------------------------------------------------------------- context and spectrum
struct ss3d_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // twiddles for lengths 2^1..2^kMaxLogN: entry k of length n at offset (n/2 - 1) + k
    DevBuf tw64f, tw64i, tw32i;
    // FP32 W_n tables for n = 2^0..2^kMaxLogN: entry m = exp(+2 pi i m/n) at offset (n - 1) + m
    DevBuf tw32n;
    Lane main;               // the context's own lane (main.stream is `stream`)
    // ss3d_realize_device_noalloc: the device address of `scratch` on this context's device, and
    // a lane on the context's streams whose buffers are windows of it
    void* scratch = nullptr;
    Lane sl;
    bool pipeline = false;   // SS3D_OPT_PIPELINE: FP32 fill of batch k+1 beside the passes of batch k
    bool fuse = true;        // SS3D_OPT_FUSE: fill of batch k+1 and y pass of batch k in one launch
    bool batch_tune = true;  // SS3D_OPT_BATCH_AUTOTUNE: batch creation autotunes every distinct grid
    bool zrun = true;        // SS3D_OPT_ZRUN: z-run layout of the FP32 fill and y outputs where it applies
    bool spec32 = false;     // SS3D_OPT_SPEC_FP32: ss3d_spectrum_create builds FP32-only spectra
    bool fp64_generic = false;  // SS3D_OPT_FP64_GENERIC: FP64 realizations through the generic pass kernel
    int target_e64 = 4096;   // elements per FP64 tile
    int target_e32 = 8192;   // elements per FP32 tile
    // per-pass event timing (ss3d_context_profile)
    struct ProfRec {
        int kind;
        hipEvent_t e0, e1;
        double bytes;
    };
    bool profiling = false;
    std::vector<ProfRec> prof;
    // autotuned radix plan of the FP32 passes: (profiling kind, log2 length) -> largest radix log
    std::map<std::pair<int, int>, int> tune;
    std::vector<hipEvent_t> event_pool;
    DevBuf vario_in, vario_out, vario_np, vario_part;  // ss3d_variogram work buffers
    // spectra created by ss3d_spectrum_create and not destroyed yet: detached (ctx = null) when the
    // context is destroyed, so that destroying them afterwards touches no freed context
    std::set<ss3d_spectrum*> spectra;
    ~ss3d_context() {
        for (auto& r : prof) {
            (void)hipEventDestroy(r.e0);
            (void)hipEventDestroy(r.e1);
        }
        for (auto e : event_pool) (void)hipEventDestroy(e);
    }
};

struct RngTables {
    bool ready = false;
    DevBuf rowf, txp;
};

struct ss3d_spectrum {
    ss3d_context* ctx = nullptr;  // null once the context is destroyed
    int device = 0;
    int npx = 0, npy = 0, npz = 0, hx = 0, hxa = 0;
    DevBuf amp64, amp32h, diagd;
    RngTables rng[2];
    bool has64 = true;          // amp64 written (the request loop builds FP32-only spectra without it)
    BufPool* pool = nullptr;    // request-loop spectra: buffers come from (and return to) this pool
    ss3d_spec_diag diag{};
    // the model this spectrum was made for (ss3d_spectrum_create / load): the key of its file
    int key_nst = 0;
    ss3d_structure key_st[kMaxStruct] = {};
    double key_nugget = 0, key_d[3] = {0, 0, 0};
};

namespace {

const void* tw_ptr(const DevBuf& b, int logn, size_t esz) {
    size_t off = logn >= 1 ? ((size_t(1) << (logn - 1)) - 1) : 0;
    return static_cast<const char*>(b.p) + off * esz;
}

int build_twiddles(ss3d_context* c) {
    const size_t total = (size_t(1) << kMaxLogN) - 1;  // sum of n/2 over n = 2..2^kMaxLogN
    std::vector<double> f(2 * total), iv(2 * total);
    std::vector<float> i32(2 * total);
    for (int L = 1; L <= kMaxLogN; ++L) {
        const int n = 1 << L;
        const size_t off = (size_t(1) << (L - 1)) - 1;
        for (int k = 0; k < n / 2; ++k) {
            // f90:249: cmplx(cos(2.0d0*pi*k/n), isign*sin(2.0d0*pi*k/n))
            const double ang = 2.0 * kPi * (double)k / (double)n;
            const double cs = std::cos(ang), sn = std::sin(ang);
            f[2 * (off + k)] = cs;
            f[2 * (off + k) + 1] = -1.0 * sn;
            iv[2 * (off + k)] = cs;
            iv[2 * (off + k) + 1] = 1.0 * sn;
            i32[2 * (off + k)] = (float)cs;
            i32[2 * (off + k) + 1] = (float)sn;
        }
    }
    std::vector<float> wn(2 * ((size_t(2) << kMaxLogN) - 1));
    for (int L = 0; L <= kMaxLogN; ++L) {
        const int n = 1 << L;
        const size_t off = (size_t(1) << L) - 1;
        for (int m = 0; m < n; ++m) {
            const double ang = 2.0 * kPi * (double)m / (double)n;
            wn[2 * (off + m)] = (float)std::cos(ang);
            wn[2 * (off + m) + 1] = (float)std::sin(ang);
        }
    }
    SS3D_TRY(c->tw32n.ensure(wn.size() * 4, c->stream));
    HIP_TRY(hipMemcpy(c->tw32n.p, wn.data(), wn.size() * 4, hipMemcpyHostToDevice));
    SS3D_TRY(c->tw64f.ensure(f.size() * 8, c->stream));
    SS3D_TRY(c->tw64i.ensure(iv.size() * 8, c->stream));
    SS3D_TRY(c->tw32i.ensure(i32.size() * 4, c->stream));
    HIP_TRY(hipMemcpy(c->tw64f.p, f.data(), f.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->tw64i.p, iv.data(), iv.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->tw32i.p, i32.data(), i32.size() * 4, hipMemcpyHostToDevice));
    return SS3D_OK;
}

// Random-stream tables of a grid (closed form of the fill walk f90:1305-1334):
//   txp[i]  = a^(g*i)
//   rowf[r] = a^(g*q0(r)), q0 = index among drawing points of the first point of drawing row r.
// Drawing planes: k = 0 (self-conjugate), 0 < k < npz/2 (forward), k = npz/2 (self-conjugate).
// In a self-conjugate plane, rows j = 0 and j = npy/2 draw for i <= npx/2, rows 0 < j < npy/2
// for every i; in a forward plane every row draws for every i.
int build_rng(ss3d_spectrum* sp, int igauss) {
    RngTables& t = sp->rng[igauss];
    if (t.ready) return SS3D_OK;
    const int npx = sp->npx, npy = sp->npy, npz = sp->npz;
    const uint64_t g = igauss ? 2 : 1;
    // tx has at least 2 entries: the FP32 fill reads the factor of column pairs (c0, c0+1)
    std::vector<uint32_t> tx(std::max(npx, 2), 0u), rowf((size_t)npy * npz, 0u);
    const uint32_t ax = pm_pow(kPmA, g);
    tx[0] = 1;
    for (int i = 1; i < npx; ++i) tx[i] = pm_mul(tx[i - 1], ax);
    const int64_t hxc = npx >= 2 ? npx / 2 + 1 : 1;
    const int64_t hp = npy >= 2 ? (int64_t)(npy / 2 - 1) * npx + 2 * hxc : hxc;
    const int64_t fp = (int64_t)npy * npx;
    const uint32_t arow = pm_pow(kPmA, g * (uint64_t)npx);
    for (int k = 0; k <= npz / 2 && k < npz; ++k) {
        const int64_t ps = (k == 0) ? 0 : hp + (int64_t)(k - 1) * fp;
        const bool self_plane = (k == 0) || (npz >= 2 && k == npz / 2);
        uint32_t* rk = rowf.data() + (size_t)npy * k;
        if (!self_plane) {
            uint32_t v = pm_pow(kPmA, g * (uint64_t)ps);
            for (int j = 0; j < npy; ++j) {
                rk[j] = v;
                v = pm_mul(v, arow);
            }
        } else {
            rk[0] = pm_pow(kPmA, g * (uint64_t)ps);
            if (npy >= 2) {
                uint32_t v = pm_pow(kPmA, g * (uint64_t)(ps + hxc));
                for (int j = 1; j <= npy / 2; ++j) {
                    rk[j] = v;
                    v = pm_mul(v, arow);
                }
            }
        }
    }
    hipStream_t s = sp->ctx->stream;
    SS3D_TRY(grab(sp->pool, t.txp, tx.size() * 4, s));
    SS3D_TRY(grab(sp->pool, t.rowf, rowf.size() * 4, s));
    HIP_TRY(hipMemcpy(t.txp.p, tx.data(), tx.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.rowf.p, rowf.data(), rowf.size() * 4, hipMemcpyHostToDevice));
    t.ready = true;
    return SS3D_OK;
}

// Tile shape of a pass: returns threads per block and sets logp.
struct Shape {
    int nt, logp;
};
Shape contig_shape(int logn, int64_t nrows, int target_e, int emax) {
    int64_t e = std::max<int64_t>(target_e, int64_t(1) << logn);
    // do not make tiles much larger than the work
    int64_t work = (int64_t)nextpow2(std::max<int64_t>(nrows, 1)) << logn;
    e = std::min(e, std::max<int64_t>(work, int64_t(1) << logn));
    e = std::max<int64_t>(e, 64 * kEPT);
    e = std::min<int64_t>(e, emax);
    return {(int)(e / kEPT), ilog2(e) - logn};
}
Shape strided_shape(int logn, int64_t n0, int row_bytes_target, int esz, int target_e, int emax) {
    int64_t tx = std::max<int64_t>(1, row_bytes_target / esz);
    tx = std::min<int64_t>(tx, nextpow2(std::max<int64_t>(n0, 1)));
    int64_t e = tx << logn;
    if (e < target_e) e = std::min<int64_t>(target_e, nextpow2(std::max<int64_t>(n0, 1)) << logn);
    e = std::max<int64_t>(e, 64 * kEPT);
    e = std::min<int64_t>(e, emax);
    return {(int)(e / kEPT), ilog2(e) - logn};
}

struct Pass {
    int prec;  // 64 or 32
    int in, out;
    Shape sh;
    int nblocks;
    PassArgs a;
    int kind;      // profiling kind (see ss3d_context_profile)
    int fast = 0;  // FP32 fast kernels: 1 strided, 2 C2R x pass, 3 one-point x axis;
                   // FP64 compile-time kernels: 4 x pass (fill + rows), 5 strided,
                   // 6 x pass of the spectral density (covariance rows)
    int rm = 5;    // FP32 fast kernels: largest radix 2^rm
    int var = 0;   // FP32 strided kernels: tile variant (fast_strided_tile / fast_strided_pad)
};

// Bytes a pass must move through HBM per batch member: the operands it reads (complex input,
// or the amplitude it fills from) and the results it writes.
// Algorithmic HBM bytes of one launch over nbatch members: outputs written once per member,
// operands read once per member, except the amplitudes of fill passes, read once per block group
// (`groups` = ceil(nbatch / members per block)).
double pass_bytes(const Pass& p, int nbatch, int groups) {
    const PassArgs& a = p.a;
    const double ec = p.prec == 64 ? 16.0 : 8.0, er = ec / 2;
    const double n = double(int64_t(1) << a.logn);
    const double cols = a.axis == 0 ? double(a.nrows) : double(a.n0) * a.nthird;
    if (p.fast == 2 || p.fast == 3)  // C2R x pass: half rows (npx/2 complex) in, real cropped rows out
        return ((p.in == IN_LOADH || p.fast == 3 ? cols * std::max(1, a.npx / 2) * ec : a.hx * 4.0) +
                cols * a.n_out * er) * nbatch;
    double in = 0, out = 0;
    switch (p.in) {
        case IN_LOAD: in = cols * n * ec; break;
        case IN_LOADH: in = cols * a.hx * ec; break;
        case IN_FILL64: in = cols * n * 8.0; break;
        case IN_FILL32H: in = cols * n * 4.0; break;
        case IN_FILLH: in = a.hx * 4.0; break;
        default: break;
    }
    switch (p.out) {
        case OUT_STORE: out = cols * a.n_out * ec; break;
        case OUT_REAL: out = cols * a.n_out * er; break;
        case OUT_SPEC: out = double(a.npx) * a.npy * a.npz * 8.0 + double(a.hx) * a.npy * a.npz * 4.0; break;
        default: break;
    }
    return in * (p.in == IN_FILL32H ? groups : nbatch) + out * nbatch;
}

hipEvent_t take_event(ss3d_context* c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Tiles, blocks and member groups of a launch of p over nbatch members (set in p.a), the
// statistics buffer of passes that write statistics; `groups` = block groups along members.
int prep_pass(Pass& p, int nbatch, hipStream_t strm, DevBuf& partial, int& groups) {
    p.a.logp = p.sh.logp;
    int P = 1 << p.sh.logp;
    if (p.fast == 1) P = fast_strided_tile(p.a.logn, p.var);
    if (p.fast == 2) P = fast_c2r_rows(p.a.logn, p.rm);
    if (p.fast == 3) P = 256;
    if (p.fast >= 4) P = fast64_tile(p.fast, p.a.logn);
    if (p.fast == 4) {
        p.nblocks = fast64_rows_blocks(p.a.logn, p.a.npy, p.a.npz);
    } else if (p.a.axis == 0) {
        p.nblocks = (int)((p.a.nrows + P - 1) / P);
    } else {
        p.a.ntiles0 = (p.a.n0 + P - 1) / P;
        p.nblocks = p.a.ntiles0 * p.a.nthird;
    }
    p.a.nblocks = p.nblocks;
    if (p.nblocks <= 0) return fail(SS3D_EINVAL, "empty pass");
    // FP32 strided passes and the C2R pass from the y output: a block serves bpb batch members, as
    // many as keep >= kFillBlocks blocks
    p.a.nbatch = nbatch;
    p.a.bpb = 1;
    if (p.fast == 1 || (p.fast == 2 && p.in == IN_LOADH))
        p.a.bpb = std::max(1, std::min(nbatch, (int)((int64_t)p.nblocks * nbatch / kFillBlocks)));
    groups = (nbatch + p.a.bpb - 1) / p.a.bpb;
    if (p.fast == 1 && p.in == IN_LOAD) {  // buffer addressing: byte offsets inside a tile below 2^31
        const int64_t span_in = ((int64_t(1) << p.a.logn) * p.a.in_sA + P) * 8;
        const int64_t span_out = ((int64_t)p.a.n_out * p.a.out_sA + P) * 8;
        if (span_in > 0x7fffffffLL || span_out > 0x7fffffffLL)
            return fail(SS3D_ETOOBIG, "FP32 pass tile spans more than 2 GiB");
    }
    if (p.out == OUT_SPEC || p.out == OUT_REAL) {
        SS3D_TRY(partial.ensure((size_t)p.nblocks * nbatch * 4 * sizeof(double), strm));
        p.a.partial = partial.as<double>();
    }
    return SS3D_OK;
}

int run_pass(ss3d_context* c, Pass& p, int nbatch, hipStream_t strm, DevBuf& partial) {
    int groups = 1;
    SS3D_TRY(prep_pass(p, nbatch, strm, partial, groups));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->profiling) {
        e0 = take_event(c);
        e1 = take_event(c);
        if (!e0 || !e1) return fail(SS3D_EHIP, "hipEventCreate failed");
        HIP_TRY(hipEventRecord(e0, strm));
    }
    hipError_t e;
    if (p.fast == 1)
        e = launch_fast_strided(p.a.logn, p.in, p.a.axis, p.rm, p.var, p.nblocks, groups, p.a, strm);
    else if (p.fast == 2)
        e = launch_fast_c2r(p.a.logn, p.in, p.rm, p.nblocks, nbatch, p.a, strm);
    else if (p.fast == 3)
        e = launch_fast_x1(p.nblocks, nbatch, p.a, strm);
    else if (p.fast >= 4)
        e = launch_fast64(p.fast, p.a.logn, p.out, p.nblocks, nbatch, p.a, strm);
    else
        e = launch_pass64(p.in, p.out, p.sh.nt, p.nblocks, nbatch, p.a, strm);
    if (e != hipSuccess) return fail(SS3D_EHIP, "pass launch (prec %d in %d out %d nt %d): %s", p.prec, p.in, p.out,
                                     p.sh.nt, hipGetErrorString(e));
    if (c->profiling) {
        HIP_TRY(hipEventRecord(e1, strm));
        c->prof.push_back({p.kind, e0, e1, pass_bytes(p, nbatch, groups)});
    }
    return SS3D_OK;
}

// The fill pass pz over nbz members beside the y pass py over nby members, in one launch of
// the fused kernel (pass kind 9).  Block roles interleave in the ratio of their block counts.
int run_fused(ss3d_context* c, Pass& pz, int nbz, Pass& py, int nby, hipStream_t strm, DevBuf& partial) {
    int gz = 1, gy = 1;
    SS3D_TRY(prep_pass(pz, nbz, strm, partial, gz));
    SS3D_TRY(prep_pass(py, nby, strm, partial, gy));
    const int64_t nzb = (int64_t)pz.nblocks * gz, nyb = (int64_t)py.nblocks * gy;
    if (nzb + nyb > 0x7fffffffLL) return fail(SS3D_ETOOBIG, "fused pass of %lld blocks", (long long)(nzb + nyb));
    FuseMap m;
    m.nzb = (int)nzb;
    m.nyb = (int)nyb;
    m.A = nzb >= nyb ? (int)std::max<int64_t>(1, (nzb + nyb / 2) / nyb) : 1;
    m.C = nzb >= nyb ? 1 : (int)std::max<int64_t>(1, (nyb + nzb / 2) / nzb);
    m.K = std::min(m.nzb / (8 * m.A), m.nyb / (8 * m.C));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->profiling) {
        e0 = take_event(c);
        e1 = take_event(c);
        if (!e0 || !e1) return fail(SS3D_EHIP, "hipEventCreate failed");
        HIP_TRY(hipEventRecord(e0, strm));
    }
    hipError_t e = launch_fast_fused(pz.a.logn, pz.rm, py.a.logn, pz.a, py.a, m, strm);
    if (e != hipSuccess) return fail(SS3D_EHIP, "fused pass launch: %s", hipGetErrorString(e));
    if (c->profiling) {
        HIP_TRY(hipEventRecord(e1, strm));
        c->prof.push_back({9, e0, e1, pass_bytes(pz, nbz, gz) + pass_bytes(py, nby, gy)});
    }
    return SS3D_OK;
}

PassArgs base_args(const ss3d_spectrum* sp) {
    PassArgs a;
    std::memset(&a, 0, sizeof a);
    a.npx = sp->npx;
    a.npy = sp->npy;
    a.npz = sp->npz;
    a.hx = sp->hx;
    a.hxa = sp->hxa;
    return a;
}

int check_grid(int npx, int npy, int npz) {
    for (int v : {npx, npy, npz})
        if (v < 1 || (v & (v - 1)) != 0) return fail(SS3D_EINVAL, "periodic grid sizes must be powers of two (got %d)", v);
    return SS3D_OK;
}

}  // namespace

// ================================================================ C ABI
namespace {
// Device address of `scratch` on `device` (current), looked up once per device: each GPU holds
// its own copy of the array at its own address.
int scratch_address(int device, void** out) {
    static std::mutex mu;
    static std::map<int, void*> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(device);
    if (it == addr.end()) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(SS3D_EHIP, "hipGetSymbolAddress returned a null scratch address");
        it = addr.emplace(device, p).first;
    }
    *out = it->second;
    return SS3D_OK;
}
}  // namespace

extern "C" {

int ss3d_version(void) { return SS3D_VERSION; }

int ss3d_last_error(char* buf, int64_t len) {
    if (!buf || len <= 0) return SS3D_EINVAL;
    std::snprintf(buf, (size_t)len, "%s", g_last_error.c_str());
    return SS3D_OK;
}

int ss3d_context_create(int device, ss3d_context** out) {
    if (!out) return fail(SS3D_EINVAL, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(SS3D_ENODEV, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(SS3D_EINVAL, "device %d out of range (0..%d)", device, ndev - 1);
    HIP_TRY(hipSetDevice(device));
    std::unique_ptr<ss3d_context> c(new ss3d_context);
    c->device = device;
    SS3D_TRY(c->main.init());
    c->stream = c->main.stream;
    c->sl.stream = c->main.stream;
    c->sl.stream2 = c->main.stream2;
    SS3D_TRY(scratch_address(device, &c->scratch));
    SS3D_TRY(build_twiddles(c.get()));
    *out = c.release();
    return SS3D_OK;
}

int ss3d_context_destroy(ss3d_context* ctx) {
    if (!ctx) return SS3D_OK;
    (void)hipSetDevice(ctx->device);
    ctx->main.sync();
    for (ss3d_spectrum* sp : ctx->spectra) sp->ctx = nullptr;
    hipStream_t s = ctx->main.stream, s2 = ctx->main.stream2;
    delete ctx;
    if (s) (void)hipStreamDestroy(s);
    if (s2) (void)hipStreamDestroy(s2);
    (void)hipGetLastError();  // results of the clean-up calls above are not reported
    return SS3D_OK;
}

int ss3d_context_synchronize(ss3d_context* ctx) {
    if (!ctx) return fail(SS3D_EINVAL, "null context");
    HIP_TRY(hipStreamSynchronize(ctx->main.stream));
    HIP_TRY(hipStreamSynchronize(ctx->main.stream2));
    return SS3D_OK;
}

int ss3d_periodic_grid(const ss3d_structure* st, int nst, double dx, double dy, double dz, int nx, int ny, int nz,
                       int ipad, int* npx, int* npy, int* npz) {
    SS3D_TRY(check_structures(st, nst));
    if (nx < 1 || ny < 1 || nz < 1) return fail(SS3D_EINVAL, "grid sizes must be positive");
    if (!npx || !npy || !npz) return fail(SS3D_EINVAL, "null output pointer");
    int64_t px = nextpow2(nx), py = nextpow2(ny), pz = nextpow2(nz);
    if (ipad == 1) {
        // reach of the covariance along each axis in cells (f90:1236-1245)
        double ex = 0, ey = 0, ez = 0;
        for (int s = 0; s < nst; ++s) {
            double rot[9], axes[9];
            setrot(st[s], rot, axes);
            const double co = cutoff(st[s].itype);
            ex = std::max(ex, co * std::sqrt(axes[0] * axes[0] + axes[3] * axes[3] + axes[6] * axes[6]) / dx);
            ey = std::max(ey, co * std::sqrt(axes[1] * axes[1] + axes[4] * axes[4] + axes[7] * axes[7]) / dy);
            ez = std::max(ez, co * std::sqrt(axes[2] * axes[2] + axes[5] * axes[5] + axes[8] * axes[8]) / dz);
        }
        const int64_t rx = (int64_t)std::ceil(ex), ry = (int64_t)std::ceil(ey), rz = (int64_t)std::ceil(ez);
        px = nextpow2(std::max<int64_t>(nx + rx, 2 * rx));
        py = nextpow2(std::max<int64_t>(ny + ry, 2 * ry));
        pz = nz > 1 ? nextpow2(std::max<int64_t>(nz + rz, 2 * rz)) : 1;
    } else if (ipad != 0) {
        return fail(SS3D_EINVAL, "ipad must be 0 or 1");
    }
    if (px > (1 << 30) || py > (1 << 30) || pz > (1 << 30)) return fail(SS3D_ETOOBIG, "periodic grid axis too long");
    *npx = (int)px;
    *npy = (int)py;
    *npz = (int)pz;
    return SS3D_OK;
}

int64_t ss3d_draws_per_realization(int npx, int npy, int npz, int igauss) {
    return draws_per_realization(npx, npy, npz, igauss);
}

int64_t ss3d_realization_state(int iseed, int npx, int npy, int npz, int igauss, int64_t r) {
    const uint64_t d = (uint64_t)draws_per_realization(npx, npy, npz, igauss) % (kPmM - 1);
    const uint64_t e = (((uint64_t)r % (kPmM - 1)) * d) % (kPmM - 1);
    return pm_mul(seed_state(iseed), pm_pow(kPmA, e));
}

// Enqueues the whole spectrum build on lane L (no host synchronisation); the diagnostics land in
// sp->diagd on the device.
// pool: buffers from (request loop) or nullptr; with64 = false skips the FP64 amplitudes (a
// spectrum for FP32 realizations only).
static int spectrum_build(ss3d_context* c, Lane& L, const ss3d_structure* st, int nst, double nugget, double dx,
                          double dy, double dz, int npx, int npy, int npz, std::unique_ptr<ss3d_spectrum>& sp,
                          BufPool* pool = nullptr, bool with64 = true) {
    SS3D_TRY(check_structures(st, nst));
    SS3D_TRY(check_grid(npx, npy, npz));
    const int lx = ilog2(npx), ly = ilog2(npy), lz = ilog2(npz);
    if (std::max({lx, ly, lz}) > 13)
        return fail(SS3D_ETOOBIG, "periodic axis longer than 8192 is not supported by the FP64 spectrum passes");
    HIP_TRY(hipSetDevice(c->device));
    sp.reset(new ss3d_spectrum);
    sp->ctx = c;
    sp->device = c->device;
    sp->npx = npx;
    sp->npy = npy;
    sp->npz = npz;
    sp->hx = npx >= 2 ? npx / 2 + 1 : 1;
    sp->hxa = (sp->hx + 1) & ~1;
    const int64_t ntot = (int64_t)npx * npy * npz;
    const int64_t nh = (int64_t)sp->hxa * npy * npz;
    sp->pool = pool;
    sp->has64 = with64;
    if (with64) SS3D_TRY(grab(pool, sp->amp64, ntot * sizeof(double), L.stream));
    SS3D_TRY(grab(pool, sp->amp32h, nh * sizeof(float), L.stream));
    HIP_TRY(hipMemsetAsync(sp->amp32h.p, 0, nh * sizeof(float), L.stream));  // padding column stays 0
    SS3D_TRY(L.wa.ensure(ntot * 2 * sizeof(double), L.stream));
    SS3D_TRY(grab(pool, sp->diagd, 4 * sizeof(double), L.stream));

    PassArgs a = base_args(sp.get());
    a.nst = nst;
    a.nugget = nugget;
    a.dx = dx;
    a.dy = dy;
    a.dz = dz;
    double c0 = nugget;  // C(0) = covariance at lag 0 (f90:1270)
    for (int s = 0; s < nst; ++s) {
        double rot[9], axes[9];
        setrot(st[s], rot, axes);
        a.st[s].itype = st[s].itype;
        a.st[s].sill = st[s].sill;
        a.st[s].ahmax = st[s].ahmax;
        std::memcpy(a.st[s].rot, rot, sizeof rot);
        c0 = c0 + st[s].sill * 1.0;
    }
    a.amp64_out = with64 ? sp->amp64.as<double>() : nullptr;
    a.amp32h_out = sp->amp32h.as<float>();
    const int emax = 8192;
    const bool has_y = npy > 1, has_z = npz > 1;
    // x pass: covariance generated in registers, transformed along x (isign = -1)
    Pass px{64, IN_COV, (has_y || has_z) ? OUT_STORE : OUT_SPEC, contig_shape(lx, (int64_t)npy * npz, c->target_e64, emax), 0, a, 0};
    px.a.axis = 0;
    px.a.logn = lx;
    px.a.nrows = npy * npz;
    px.a.n_out = npx;
    px.a.out = L.wa.p;
    px.a.out_row = npx;
    px.a.tw = tw_ptr(c->tw64f, lx, 16);
    // compile-time kernels (same arithmetic) where the lengths allow, as for realizations
    const bool f64ok = !c->fp64_generic;
    auto fits = [](int l) { return l >= 1 && l <= kMaxFast64Log; };
    if (f64ok && (has_y || has_z) && fits(lx)) px.fast = 6;
    // FP32-only spectra: the density of a real covariance grid is Hermitian, S(-k) = conj S(k), and the
    // FP32 amplitudes hold the columns kx <= npx/2, so the passes after the x pass transform those
    // columns alone (each column by the same operations as in a full build) and the OUT_SPEC sums count
    // the columns 0 < kx < npx/2 twice (Re S(-k) = Re S(k))
    const bool half = !with64 && px.fast == 6 && (!has_y || fits(ly)) && (!has_z || fits(lz)) && sp->hx < npx;
    const int ncol = half ? sp->hx : npx;
    // row stride of W: the kept columns rounded up to whole 128-byte column tiles (z planes of about half
    // the size for the z pass to walk), at most npx (W holds npx npy npz values)
    const int64_t ws = half ? std::min((ncol + 7) & ~7, npx) : npx;
    if (half) px.a.n_out = ncol;
    px.a.out_row = ws;
    SS3D_TRY(run_pass(c, px, 1, L.stream, L.partial));
    Pass* last = &px;
    Pass py{64, IN_LOAD, has_z ? OUT_STORE : OUT_SPEC, strided_shape(ly, npx, 128, 16, c->target_e64, emax), 0, a, 1};
    if (has_y) {
        py.a.axis = 1;
        py.a.logn = ly;
        py.a.n0 = ncol;
        py.a.nthird = npz;
        py.a.n_out = npy;
        py.a.spec_half = half && !has_z;
        py.a.in = L.wa.p;
        py.a.out = L.wa.p;
        py.a.in_sA = py.a.out_sA = ws;
        py.a.in_sT = py.a.out_sT = ws * npy;
        py.a.tw = tw_ptr(c->tw64f, ly, 16);
        if (f64ok && fits(ly)) py.fast = 5;
        SS3D_TRY(run_pass(c, py, 1, L.stream, L.partial));
        last = &py;
    }
    Pass pz{64, IN_LOAD, OUT_SPEC, strided_shape(lz, npx, 128, 16, c->target_e64, emax), 0, a, 2};
    if (has_z) {
        pz.a.axis = 2;
        pz.a.logn = lz;
        pz.a.n0 = ncol;
        pz.a.nthird = npy;
        pz.a.n_out = npz;
        pz.a.spec_half = half;
        pz.a.in = L.wa.p;
        pz.a.out = L.wa.p;
        pz.a.in_sA = pz.a.out_sA = ws * npy;
        pz.a.in_sT = pz.a.out_sT = ws;
        pz.a.tw = tw_ptr(c->tw64f, lz, 16);
        if (f64ok && fits(lz)) pz.fast = 5;
        SS3D_TRY(run_pass(c, pz, 1, L.stream, L.partial));
        last = &pz;
    }
    hipError_t e = launch_diag_finalize(L.partial.as<double>(), last->nblocks, sp->diagd.as<double>(), L.stream);
    if (e != hipSuccess) return fail(SS3D_EHIP, "diag finalize: %s", hipGetErrorString(e));
    sp->diag.c0 = c0;
    return SS3D_OK;
}

// Copies the device diagnostics of a finished build into sp->diag (synchronous).
static int spectrum_read_diag(ss3d_spectrum* sp, hipStream_t s) {
    double d4[4];
    HIP_TRY(hipMemcpyAsync(d4, sp->diagd.p, sizeof d4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    sp->diag.max_imag = d4[0];
    sp->diag.max_real = d4[1];
    sp->diag.sneg = d4[2];
    sp->diag.spos = d4[3];
    return SS3D_OK;
}

int ss3d_spectrum_create(ss3d_context* c, const ss3d_structure* st, int nst, double nugget, double dx, double dy,
                         double dz, int npx, int npy, int npz, ss3d_spectrum** out, ss3d_spec_diag* diag) {
    if (!c || !out) return fail(SS3D_EINVAL, "null argument");
    *out = nullptr;
    std::unique_ptr<ss3d_spectrum> sp;
    SS3D_TRY(spectrum_build(c, c->main, st, nst, nugget, dx, dy, dz, npx, npy, npz, sp, nullptr, !c->spec32));
    SS3D_TRY(spectrum_read_diag(sp.get(), c->main.stream));
    if (diag) *diag = sp->diag;
    sp->key_nst = nst;
    std::memcpy(sp->key_st, st, sizeof(ss3d_structure) * nst);
    sp->key_nugget = nugget;
    sp->key_d[0] = dx;
    sp->key_d[1] = dy;
    sp->key_d[2] = dz;
    c->spectra.insert(sp.get());
    *out = sp.release();
    return SS3D_OK;
}

int ss3d_context_profile(ss3d_context* c, int enable) {
    if (!c) return fail(SS3D_EINVAL, "null context");
    if (enable) {
        c->main.sync();
        for (auto& r : c->prof) {
            c->event_pool.push_back(r.e0);
            c->event_pool.push_back(r.e1);
        }
        c->prof.clear();
    }
    c->profiling = enable != 0;
    return SS3D_OK;
}

int ss3d_context_profile_read(ss3d_context* c, double* ms, double* bytes, int64_t* launches, int nkinds) {
    if (!c || !ms || !bytes || !launches || nkinds < 1) return fail(SS3D_EINVAL, "null argument");
    c->main.sync();
    for (int k = 0; k < nkinds; ++k) {
        ms[k] = 0;
        bytes[k] = 0;
        launches[k] = 0;
    }
    for (auto& r : c->prof) {
        if (r.kind < 0 || r.kind >= nkinds) continue;
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, r.e0, r.e1));
        ms[r.kind] += t;
        bytes[r.kind] += r.bytes;
        launches[r.kind] += 1;
    }
    return SS3D_OK;
}

int ss3d_spectrum_destroy(ss3d_spectrum* sp) {
    if (!sp) return SS3D_OK;
    (void)hipSetDevice(sp->device);
    if (sp->ctx) {  // else its context is gone, and its streams were drained when it was destroyed
        (void)hipStreamSynchronize(sp->ctx->stream);
        sp->ctx->spectra.erase(sp);
    }
    delete sp;
    (void)hipGetLastError();  // results of the clean-up calls above are not reported
    return SS3D_OK;
}

int64_t ss3d_spectrum_bytes(const ss3d_spectrum* sp) {
    if (!sp) return 0;
    return (int64_t)(sp->amp64.cap + sp->amp32h.cap + sp->rng[0].rowf.cap + sp->rng[0].txp.cap + sp->rng[1].rowf.cap +
                     sp->rng[1].txp.cap);
}

int ss3d_spectrum_amplitude(const ss3d_spectrum* sp, double* host_amp) {
    if (!sp || !host_amp) return fail(SS3D_EINVAL, "null argument");
    if (!sp->ctx) return fail(SS3D_EINVAL, "the spectrum's context was destroyed");
    HIP_TRY(hipSetDevice(sp->ctx->device));
    HIP_TRY(hipMemcpyAsync(host_amp, sp->amp64.p, (size_t)sp->npx * sp->npy * sp->npz * sizeof(double),
                           hipMemcpyDeviceToHost, sp->ctx->stream));
    HIP_TRY(hipStreamSynchronize(sp->ctx->stream));
    return SS3D_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- realizations
namespace {

int batch_size(const ss3d_spectrum* sp, int nx, int ny, int nz, int prec, int nreal);

// The tuning table holds rm + 16 var: the radix plan and the tile variant of a pass kind and length.
int tuned_rm(const ss3d_context* c, int kind, int L) {
    auto it = c->tune.find({kind, L});
    return it == c->tune.end() ? 5 : it->second & 15;
}
int tuned_var(const ss3d_context* c, int kind, int L) {
    auto it = c->tune.find({kind, L});
    return it == c->tune.end() ? 0 : it->second >> 4;
}

// Columns per tile of the FP32 fill output when it is laid out in column tiles (the z and y passes
// have one length and one tile variant, so a y tile reads one contiguous slab), else 0.
int fill_column_tile(const ss3d_context* c, const ss3d_spectrum* sp) {
    const int ly = ilog2(sp->npy), lz = ilog2(sp->npz);
    if (sp->npy < 2 || sp->npz < 2 || ly != lz || tuned_var(c, 6, lz) != tuned_var(c, 7, ly)) return 0;
    return fast_strided_tile(lz, tuned_var(c, 6, lz));
}

// Enqueue realizations r0..r0+nb-1 into dout (device) and their statistics into dstats.
int enqueue_batch(ss3d_context* c, Lane& L, ss3d_spectrum* sp, int iseed, int igauss, int64_t r0, int nb, int nx, int ny,
                  int nz, int prec, void* dout, double* dstats) {
    const int npx = sp->npx, npy = sp->npy, npz = sp->npz;
    const int lx = ilog2(npx), ly = ilog2(npy), lz = ilog2(npz);
    if (!sp->has64) return fail(SS3D_EINVAL, "spectrum built without FP64 amplitudes");
    SS3D_TRY(build_rng(sp, igauss));
    SS3D_TRY(L.rbase.ensure((size_t)nb * 4, L.stream));
    {
        const uint64_t dmod = (uint64_t)draws_per_realization(npx, npy, npz, igauss) % (kPmM - 1);
        hipLaunchKernelGGL(rbase_kernel, dim3((nb + 63) / 64), dim3(64), 0, L.stream, seed_state(iseed), dmod,
                           (uint64_t)r0, nb, L.rbase.as<uint32_t>());
        HIP_TRY(hipGetLastError());
    }
    PassArgs a = base_args(sp);
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.igauss = igauss;
    a.amp64 = sp->amp64.as<double>();
    a.amp32h = sp->amp32h.as<float>();
    a.rowf = sp->rng[igauss].rowf.as<uint32_t>();
    a.txp = sp->rng[igauss].txp.as<uint32_t>();
    a.rbase = L.rbase.as<uint32_t>();
    const bool has_y = npy > 1, has_z = npz > 1;
    const int64_t ncrop = (int64_t)nx * ny * nz;
    Pass* last = nullptr;
    Pass p1, p2, p3;
    if (prec == SS3D_FP64) {
        // x -> y -> z, C2C, pruned to the crop after each pass (f90:1335-1342)
        const int emax = 8192;
        const int64_t b1 = (int64_t)nx * npy * npz, b2 = (int64_t)nx * ny * npz;
        SS3D_TRY(L.wa.ensure((size_t)b1 * nb * 16, L.stream));
        SS3D_TRY(L.wb.ensure((size_t)std::max<int64_t>(b2, 1) * nb * 16, L.stream));
        p1 = Pass{64, IN_FILL64, (has_y || has_z) ? OUT_STORE : OUT_REAL,
                  contig_shape(lx, (int64_t)npy * npz, c->target_e64, emax), 0, a, 3};
        p1.a.axis = 0;
        p1.a.logn = lx;
        p1.a.nrows = npy * npz;
        p1.a.n_out = nx;
        p1.a.tw = tw_ptr(c->tw64i, lx, 16);
        // compile-time kernels (same arithmetic) where the lengths allow
        const bool f64ok = !c->fp64_generic;
        auto fits = [](int l) { return l >= 1 && l <= kMaxFast64Log; };
        if (f64ok && (has_y || has_z) && fits(lx)) p1.fast = 4;
        if (has_y || has_z) {
            p1.a.out = L.wa.p;
            p1.a.out_row = nx;
            p1.a.out_bstride = b1;
        } else {
            p1.a.out = dout;
            p1.a.out_row = nx;
            p1.a.out_bstride = ncrop;
        }
        SS3D_TRY(run_pass(c, p1, nb, L.stream, L.partial));
        last = &p1;
        if (has_y) {
            p2 = Pass{64, IN_LOAD, has_z ? OUT_STORE : OUT_REAL, strided_shape(ly, nx, 128, 16, c->target_e64, emax), 0, a, 4};
            p2.a.axis = 1;
            p2.a.logn = ly;
            p2.a.n0 = nx;
            p2.a.nthird = npz;
            p2.a.n_out = ny;
            p2.a.in = L.wa.p;
            p2.a.in_sA = nx;
            p2.a.in_sT = (int64_t)nx * npy;
            p2.a.in_bstride = b1;
            p2.a.out = has_z ? L.wb.p : dout;
            p2.a.out_sA = nx;
            p2.a.out_sT = (int64_t)nx * ny;
            p2.a.out_bstride = has_z ? b2 : ncrop;
            p2.a.tw = tw_ptr(c->tw64i, ly, 16);
            if (f64ok && fits(ly)) p2.fast = 5;
            SS3D_TRY(run_pass(c, p2, nb, L.stream, L.partial));
            last = &p2;
        }
        if (has_z) {
            p3 = Pass{64, IN_LOAD, OUT_REAL, strided_shape(lz, nx, 128, 16, c->target_e64, emax), 0, a, 5};
            p3.a.axis = 2;
            p3.a.logn = lz;
            p3.a.n0 = nx;
            p3.a.nthird = ny;
            p3.a.n_out = nz;
            p3.a.in = has_y ? L.wb.p : L.wa.p;
            p3.a.in_sA = (int64_t)nx * ny;
            p3.a.in_sT = nx;
            p3.a.in_bstride = has_y ? b2 : b1;
            p3.a.out = dout;
            p3.a.out_sA = (int64_t)nx * ny;
            p3.a.out_sT = nx;
            p3.a.out_bstride = ncrop;
            p3.a.tw = tw_ptr(c->tw64i, lz, 16);
            if (f64ok && fits(lz)) p3.fast = 5;
            SS3D_TRY(run_pass(c, p3, nb, L.stream, L.partial));
            last = &p3;
        }
    } else {
        return fail(SS3D_EINVAL, "enqueue_batch handles the FP64 path only");
    }
    hipError_t e = launch_stats_finalize(L.partial.as<double>(), last->nblocks, nb, (double)ncrop, dstats, L.stream);
    if (e != hipSuccess) return fail(SS3D_EHIP, "stats finalize: %s", hipGetErrorString(e));
    return SS3D_OK;
}

// Grows a buffer shared by both streams (waits for both before freeing the old one).
int ensure2(Lane& L, DevBuf& b, size_t bytes) {
    if (bytes <= b.cap) return SS3D_OK;
    HIP_TRY(hipStreamSynchronize(L.stream2));
    return b.ensure(bytes, L.stream);
}

// FP32 realizations r0..r0+nreal-1: z -> y passes over the half spectrum (column 0 also carries
// the Nyquist column), then the C2R x pass, in batches of nb realizations.  The fill pass (the
// first one) runs on c->stream and writes one of two pipe buffers; the other passes and the
// statistics run on L.stream2, so the fill of batch k+1 (arithmetic-bound) overlaps the
// memory-bound passes of batch k.
// Work-buffer sizes of the FP32 passes per batch member, in complex elements: the fill pass's
// output (column-tiled when the z and y axes have one length) and the y pass's output.
struct Fp32Work {
    int64_t fill, yout;
    bool staged, has_wb;
};
Fp32Work fp32_work(const ss3d_spectrum* sp, int nx, int ny, int nz, int ctile) {
    const int npx = sp->npx, npy = sp->npy, npz = sp->npz;
    const bool has_y = npy > 1, has_z = npz > 1;
    const int hc = npx >= 2 ? npx / 2 : 1, hcs = std::max(2, hc);
    const int64_t bz = (int64_t)hcs * npy * nz, by = (int64_t)hcs * ny * nz;
    int64_t bzs = bz;
    if (ctile > 0) {
        const int64_t slab = (int64_t)npy * ctile, plane = (int64_t)((hc + ctile - 1) / ctile) * slab;
        bzs = plane * nz;
    }
    Fp32Work w;
    w.staged = has_y || has_z;
    w.has_wb = has_z && has_y;
    w.fill = std::max<int64_t>(has_z ? bzs : by, 1);
    w.yout = std::max<int64_t>(by, 1);
    return w;
}

// nb_cap > 0 bounds the realizations per launch (ss3d_realize_device_noalloc: what its scratch
// windows hold).
int enqueue_fp32(ss3d_context* c, Lane& L, ss3d_spectrum* sp, int iseed, int igauss, int64_t r0, int nreal, int nx, int ny,
                 int nz, void* dout, double* dstats, int nb_cap = 0) {
    const int npx = sp->npx, npy = sp->npy, npz = sp->npz;
    const int lx = ilog2(npx), ly = ilog2(npy), lz = ilog2(npz);
    if (lx > kMaxFastLog || ly > kMaxFastLog || lz > kMaxFastLog)
        return fail(SS3D_ETOOBIG, "periodic axis longer than %d", 1 << kMaxFastLog);
    SS3D_TRY(build_rng(sp, igauss));
    SS3D_TRY(ensure2(L, L.rbase, (size_t)nreal * 4));
    {
        const uint64_t dmod = (uint64_t)draws_per_realization(npx, npy, npz, igauss) % (kPmM - 1);
        hipLaunchKernelGGL(rbase_kernel, dim3((nreal + 63) / 64), dim3(64), 0, L.stream, seed_state(iseed), dmod,
                           (uint64_t)r0, nreal, L.rbase.as<uint32_t>());
        HIP_TRY(hipGetLastError());
    }
    PassArgs a = base_args(sp);
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.igauss = igauss;
    a.amp32h = sp->amp32h.as<float>();
    a.rowf = sp->rng[igauss].rowf.as<uint32_t>();
    a.txp = sp->rng[igauss].txp.as<uint32_t>();
    const bool has_y = npy > 1, has_z = npz > 1;
    const int64_t ncrop = (int64_t)nx * ny * nz;
    const int nb = nb_cap > 0 ? std::min(nb_cap, batch_size(sp, nx, ny, nz, SS3D_FP32, nreal))
                              : batch_size(sp, nx, ny, nz, SS3D_FP32, nreal);
    const int hc = npx >= 2 ? npx / 2 : 1, hcs = std::max(2, hc);
    const int64_t bz = (int64_t)hcs * npy * nz, by = (int64_t)hcs * ny * nz;
    const float2* tw32 = c->tw32n.as<float2>();
    auto wtab = [&](int L) { return static_cast<const void*>(tw32 + ((size_t(1) << L) - 1)); };

    // pass templates; `in`/`out` buffers and rbase are set per batch
    Pass pz{32, IN_FILL32H, OUT_STORE, Shape{0, 0}, 0, a, 6, 1};
    pz.rm = tuned_rm(c, 6, lz);
    pz.var = tuned_var(c, 6, lz);
    pz.a.axis = 2;
    pz.a.logn = lz;
    pz.a.n0 = hc;
    pz.a.nthird = npy;
    pz.a.n_out = nz;
    pz.a.out_sA = (int64_t)hcs * npy;
    pz.a.out_sT = hcs;
    pz.a.out_bstride = bz;
    pz.a.tw = wtab(lz);
    Pass py{32, has_z ? IN_LOAD : IN_FILL32H, OUT_STORE, Shape{0, 0}, 0, a, 7, 1};
    py.rm = tuned_rm(c, 7, ly);
    py.var = tuned_var(c, 7, ly);
    py.a.axis = 1;
    py.a.logn = ly;
    py.a.n0 = hc;
    py.a.nthird = nz;
    py.a.n_out = ny;
    py.a.in_sA = hcs;
    py.a.in_sT = (int64_t)hcs * npy;
    py.a.in_bstride = bz;
    py.a.out_sA = hcs;
    py.a.out_sT = (int64_t)hcs * ny;
    py.a.out_bstride = by;
    py.a.tw = wtab(ly);
    // z and y of one length and tile variant (one tile width P): the fill writes BZ in column
    // tiles, a slab of P columns x npy rows per tile and z plane, so that a y tile reads one
    // contiguous slab
    int64_t bzs = bz;
    if (const int P = fill_column_tile(c, sp)) {
        const int64_t slab = (int64_t)npy * P, plane = (int64_t)((hc + P - 1) / P) * slab;
        bzs = plane * nz;
        pz.a.out_sA = plane;
        pz.a.out_sT = P;
        pz.a.out_sC = slab;
        pz.a.out_bstride = bzs;
        py.a.in_sA = P;
        py.a.in_sT = plane;
        py.a.in_sC = slab;
        py.a.in_bstride = bzs;
    }
    // x pass: npx >= 2 -> C2R on complex length npx/2; npx = 1 -> real part of each row
    Pass px{32, (has_y || has_z) ? IN_LOADH : IN_FILLH, OUT_REAL, Shape{0, 0}, 0, a, 8, npx >= 2 ? 2 : 3};
    px.rm = tuned_rm(c, 8, npx >= 2 ? lx - 1 : 0);
    px.a.axis = 0;
    px.a.logn = npx >= 2 ? lx - 1 : 0;
    px.a.nrows = ny * nz;
    px.a.n_out = nx;
    px.a.in_row = hcs;
    px.a.in_bstride = has_y ? by : bz;
    px.a.out_row = nx;
    px.a.out_bstride = ncrop;
    px.a.tw = wtab(npx >= 2 ? lx - 1 : 0);
    // z-run layout, in pairs of complex (a float4: kx even and odd) with hp = hc/2 kx pairs:
    //   fill output  [kx pair][z/8][ky][z mod 8]  (npy rows ky, the cropped nz): the runs a fill
    //                wave stores for one pair and row lie within 64 KB of each other
    //   y output     [z/8][kx pair][y][z mod 8]   (the cropped ny): the pairs of a C2R tile's
    //                rows lie 16 ny bytes apart
    // The y pass takes tiles (kx pair, z/8) of 16 columns (z mod 8, kx mod 2): a slab of 128-byte
    // runs per tile, read and written contiguously; the C2R pass reads rows (y, z) a run at a time.
    const int pc = npx >= 2 ? fast_zrun_c2r_rows(lx - 1, px.rm) : 0;
    const bool zrun = c->zrun && has_z && has_y && lz == fast_zrun_log() && ly == lz && pz.var == 0 && py.var == 0 &&
                      hc % 16 == 0 && nz % 8 == 0 && pc > 0 && ny % (pc / 8) == 0 && bz * 8 <= 0x7fffffffLL &&
                      by * 8 <= 0x7fffffffLL && ncrop * 4 <= 0x7fffffffLL;
    if (zrun) {
        const int64_t hp = hc / 2;
        pz.a.zrun = 1;
        pz.a.out_sT = 16;                   // per ky
        pz.a.out_sA = 16 * npy;             // per run of 8 z
        pz.a.out_sC = 16 * npy * (nz / 8);  // per kx pair
        pz.a.out_bstride = bz;
        py.a.n0 = (int)(16 * hp);  // column 16 p + 2 (z mod 8) + kx mod 2 of kx pair p
        py.a.nthird = nz / 8;
        py.a.in_sA = 16;
        py.a.in_sT = 16 * npy;
        py.a.in_sC = 16 * npy * (nz / 8);
        py.a.in_bstride = bz;
        py.a.out_sA = 16;
        py.a.out_sC = 16 * ny;
        py.a.out_sT = 16 * ny * hp;
        py.a.out_bstride = by;
        px.a.zrun = 1;
        px.a.in_sA = 16;              // per y
        px.a.in_sC = 16 * ny;         // per kx pair
        px.a.in_sT = 16 * ny * hp;    // per run of 8 z
    }

    const bool staged = has_y || has_z;  // a fill pass followed by further passes
    const Fp32Work wk = fp32_work(sp, nx, ny, nz, fill_column_tile(c, sp));
    hipStream_t s2 = c->pipeline ? L.stream2 : L.stream;  // stream of the passes after the fill
    const bool pipelined = staged && c->pipeline;
    if (staged) {
        for (int i = 0; i < 2; ++i) SS3D_TRY(ensure2(L, L.pipe[i], (size_t)wk.fill * nb * 8));
        if (wk.has_wb) SS3D_TRY(ensure2(L, L.wb, (size_t)wk.yout * nb * 8));
    }
    const int nbat = (nreal + nb - 1) / nb;
    // fusion keeps the y pass's own radix plan (results do not depend on the schedule): the fill's
    // plan at the same length, plan 3 at twice the length
    const int rmy_fused = lz == ly ? pz.rm : 3;
    const bool fuse2 = staged && has_z && has_y && !pipelined && c->fuse && nbat >= 2 && pz.var == 0 &&
                       py.var == 0 && fast_same_plan(ly, py.rm, rmy_fused) && fast_fusable(lz, pz.rm, ly, rmy_fused);
    if (fuse2) {
        // launch k: fill of batch k beside the y pass of batch k-1 (pipe buffers alternate),
        // then the x pass of batch k-1
        py.rm = rmy_fused;  // the same plan, the template instance of the fused kernel
        for (int k = 0; k <= nbat; ++k) {
            const bool zk = k < nbat, yk = k >= 1;
            const int cz = zk ? std::min(nb, nreal - k * nb) : 0;
            const int p0 = (k - 1) * nb, cy = yk ? std::min(nb, nreal - p0) : 0;
            if (zk) {
                pz.a.rbase = L.rbase.as<uint32_t>() + (size_t)k * nb;
                pz.a.out = L.pipe[k & 1].p;
            }
            if (yk) {
                py.a.in = L.pipe[(k - 1) & 1].p;
                py.a.out = L.wb.p;
            }
            if (zk && yk)
                SS3D_TRY(run_fused(c, pz, cz, py, cy, L.stream, L.partial));
            else if (zk)
                SS3D_TRY(run_pass(c, pz, cz, L.stream, L.partial));
            else
                SS3D_TRY(run_pass(c, py, cy, L.stream, L.partial));
            if (!yk) continue;
            px.a.in = L.wb.p;
            px.a.out = static_cast<char*>(dout) + (size_t)p0 * ncrop * 4;
            SS3D_TRY(run_pass(c, px, cy, L.stream, L.partial));
            hipError_t e = launch_stats_finalize(L.partial.as<double>(), px.nblocks, cy, (double)ncrop,
                                                 dstats + 3 * (size_t)p0, L.stream);
            if (e != hipSuccess) return fail(SS3D_EHIP, "stats finalize: %s", hipGetErrorString(e));
        }
        return SS3D_OK;
    }
    if (pipelined) {
        HIP_TRY(hipEventRecord(L.ev_join, L.stream));
        HIP_TRY(hipStreamWaitEvent(L.stream2, L.ev_join, 0));
    }
    for (int k = 0; k < nbat; ++k) {
        const int done = k * nb, cnt = std::min(nb, nreal - done);
        const int buf = k & 1;
        void* out = static_cast<char*>(dout) + (size_t)done * ncrop * 4;
        double* st = dstats + 3 * (size_t)done;
        const uint32_t* rb = L.rbase.as<uint32_t>() + done;
        if (!staged) {
            px.a.rbase = rb;
            px.a.out = out;
            SS3D_TRY(run_pass(c, px, cnt, L.stream, L.partial));
            hipError_t e = launch_stats_finalize(L.partial.as<double>(), px.nblocks, cnt, (double)ncrop, st, L.stream);
            if (e != hipSuccess) return fail(SS3D_EHIP, "stats finalize: %s", hipGetErrorString(e));
            continue;
        }
        if (pipelined && k >= 2) HIP_TRY(hipStreamWaitEvent(L.stream, L.ev_rest[buf], 0));
        Pass& fill = has_z ? pz : py;
        fill.a.rbase = rb;
        fill.a.out = L.pipe[buf].p;
        SS3D_TRY(run_pass(c, fill, cnt, L.stream, L.partial));
        if (pipelined) {
            HIP_TRY(hipEventRecord(L.ev_fill[buf], L.stream));
            HIP_TRY(hipStreamWaitEvent(L.stream2, L.ev_fill[buf], 0));
        }
        DevBuf& part = pipelined ? L.partial2 : L.partial;
        if (has_z && has_y) {
            py.a.in = L.pipe[buf].p;
            py.a.out = L.wb.p;
            SS3D_TRY(run_pass(c, py, cnt, s2, part));
            px.a.in = L.wb.p;
        } else {
            px.a.in = L.pipe[buf].p;
        }
        px.a.out = out;
        SS3D_TRY(run_pass(c, px, cnt, s2, part));
        hipError_t e = launch_stats_finalize(part.as<double>(), px.nblocks, cnt, (double)ncrop, st, s2);
        if (e != hipSuccess) return fail(SS3D_EHIP, "stats finalize: %s", hipGetErrorString(e));
        if (pipelined) HIP_TRY(hipEventRecord(L.ev_rest[buf], L.stream2));
    }
    if (pipelined) {
        HIP_TRY(hipEventRecord(L.ev_join, L.stream2));
        HIP_TRY(hipStreamWaitEvent(L.stream, L.ev_join, 0));
    }
    return SS3D_OK;
}

int check_realize(ss3d_context* c, const ss3d_spectrum* sp, int igauss, int64_t r0, int nreal, int nx, int ny, int nz,
                  int prec) {
    if (!c || !sp) return fail(SS3D_EINVAL, "null context or spectrum");
    if (sp->ctx != c) return fail(SS3D_EINVAL, "spectrum belongs to another (or a destroyed) context");
    if (igauss != 0 && igauss != 1) return fail(SS3D_EINVAL, "igauss must be 0 or 1");
    if (prec != SS3D_FP64 && prec != SS3D_FP32) return fail(SS3D_EINVAL, "precision must be SS3D_FP64 or SS3D_FP32");
    if (r0 < 0 || nreal < 0) return fail(SS3D_EINVAL, "realization range must be non-negative");
    if (nx < 1 || ny < 1 || nz < 1 || nx > sp->npx || ny > sp->npy || nz > sp->npz)
        return fail(SS3D_EINVAL, "output grid %dx%dx%d must be positive and inside the periodic grid %dx%dx%d", nx, ny, nz,
                    sp->npx, sp->npy, sp->npz);
    return SS3D_OK;
}

// realizations per launch: enough to fill the GPU for small grids, bounded by work memory
int batch_size(const ss3d_spectrum* sp, int nx, int ny, int nz, int prec, int nreal) {
    const int64_t per = prec == SS3D_FP64
                            ? ((int64_t)nx * sp->npy * sp->npz + (int64_t)nx * ny * sp->npz) * 16
                            : ((int64_t)std::max(2, sp->npx / 2) * (sp->npy * (int64_t)nz + (int64_t)ny * nz)) * 8;
    // `per` counts one fill output and one y output per member (the second fill buffer of the
    // fused launch comes on top).  16 GiB of it (of 288 GB): FP32 realizations 16 per launch at
    // 512^3 periodic grids and 10 at 1024x1024x512, whose fill pass then amortizes its per-site
    // work over that many members (at 1024x1024x512, 8 members gave 63 against 54 Gcells/s for
    // one; at 512^3, 16 members per launch run the 256^3 step 2-3% faster than 8); FP64
    // realizations 8 per launch (10 ran the 64 realizations of the 256^3 step 3% slower: 6 x 10 + 4)
    const int64_t budget = int64_t(16) << 30;
    int64_t nb = std::max<int64_t>(1, budget / std::max<int64_t>(per, 1));
    // small grids: enough realizations per launch to fill the GPU; large grids: up to 16 (FP32) or 8
    const int64_t fill = (int64_t(1) << 24) / std::max<int64_t>((int64_t)sp->npx * sp->npy * sp->npz, 1);
    nb = std::min<int64_t>(nb, std::max<int64_t>(fill, prec == SS3D_FP64 ? 8 : 16));
    nb = std::max<int64_t>(std::min<int64_t>(nb, nreal), 1);
    // the fewest batches of at most nb realizations, as even as possible (32 realizations at most
    // 10 per launch: 4 x 8 rather than 10, 10, 10, 2)
    const int64_t nbat = (std::max(nreal, 1) + nb - 1) / nb;
    return (int)((std::max(nreal, 1) + nbat - 1) / nbat);
}

}  // namespace

extern "C" {

int ss3d_realize_device(ss3d_context* c, const ss3d_spectrum* spc, int iseed, int igauss, int64_t r0, int nreal, int nx,
                        int ny, int nz, int prec, void* dev_out, double* dev_stats) {
    SS3D_TRY(check_realize(c, spc, igauss, r0, nreal, nx, ny, nz, prec));
    if (nreal == 0) return SS3D_OK;
    if (!dev_out) return fail(SS3D_EINVAL, "null device output");
    HIP_TRY(hipSetDevice(c->device));
    ss3d_spectrum* sp = const_cast<ss3d_spectrum*>(spc);
    const int64_t ncrop = (int64_t)nx * ny * nz;
    if (prec == SS3D_FP32) {
        if (!dev_stats) SS3D_TRY(ensure2(c->main, c->main.stats, (size_t)nreal * 3 * sizeof(double)));
        return enqueue_fp32(c, c->main, sp, iseed, igauss, r0, nreal, nx, ny, nz, dev_out,
                            dev_stats ? dev_stats : c->main.stats.as<double>());
    }
    const int nb = batch_size(sp, nx, ny, nz, prec, nreal);
    if (!dev_stats) SS3D_TRY(ensure2(c->main, c->main.stats, (size_t)nreal * 3 * sizeof(double)));
    for (int done = 0; done < nreal; done += nb) {
        const int cnt = std::min(nb, nreal - done);
        void* out = static_cast<char*>(dev_out) + (size_t)done * ncrop * 8;
        double* st = (dev_stats ? dev_stats : c->main.stats.as<double>()) + 3 * (size_t)done;
        SS3D_TRY(enqueue_batch(c, c->main, sp, iseed, igauss, r0 + done, cnt, nx, ny, nz, prec, out, st));
    }
    return SS3D_OK;
}

}  // extern "C"

namespace {

size_t up256(size_t b) { return (b + 255) & ~size_t(255); }

// Scratch layout of ss3d_realize_device_noalloc: realization bases and (when the caller passes
// no statistics buffer) statistics for `chunk` realizations per enqueue, then per batch member
// the block statistics and the pass buffers, for nb members per launch.
struct ScratchPlan {
    int nb = 0, chunk = 0;
    size_t rbase = 0, stats = 0, partial = 0, w1 = 0, w2 = 0, w3 = 0;
};

int plan_scratch(const ss3d_context* c, const ss3d_spectrum* sp, int nx, int ny, int nz, int prec, int nreal,
                 bool own_stats, ScratchPlan& q) {
    const int64_t npy = sp->npy, npz = sp->npz;
    int64_t blocks, per1, per2 = 0, per3 = 0;  // blocks with statistics, buffer bytes per member
    if (prec == SS3D_FP32) {
        const Fp32Work w = fp32_work(sp, nx, ny, nz, fill_column_tile(c, sp));
        blocks = (int64_t)ny * nz;  // x pass: at most one block per row
        if (w.staged) {
            per1 = per2 = w.fill * 8;
            per3 = w.has_wb ? w.yout * 8 : 0;
        } else {
            per1 = 0;
        }
    } else {
        // last pass: x rows (npy npz), y tiles (<= nx npz) or z tiles (<= nx ny)
        blocks = std::max<int64_t>(npy * npz, std::max<int64_t>((int64_t)nx * npz, (int64_t)nx * ny));
        per1 = (int64_t)nx * npy * npz * 16;
        per2 = std::max<int64_t>((int64_t)nx * ny * npz, 1) * 16;
    }
    const int64_t per = (int64_t)up256(blocks * 4 * sizeof(double)) + (int64_t)up256(per1) + (int64_t)up256(per2) +
                        (int64_t)up256(per3);
    q.chunk = std::min(nreal, 4096);
    const size_t fixed = up256((size_t)q.chunk * 4) + (own_stats ? up256((size_t)q.chunk * 3 * sizeof(double)) : 0);
    const int64_t room = (int64_t)kScratchBytes - (int64_t)fixed;
    if (room < per)
        return fail(SS3D_ETOOBIG,
                    "ss3d_realize_device_noalloc: one realization of this grid needs %lld bytes of scratch, the "
                    "scratch holds %zu (use ss3d_realize_device)",
                    (long long)(per + (int64_t)fixed), kScratchBytes);
    q.nb = (int)std::min<int64_t>({room / per, (int64_t)batch_size(sp, nx, ny, nz, prec, nreal), (int64_t)q.chunk});
    q.rbase = (size_t)q.chunk * 4;
    q.stats = own_stats ? (size_t)q.chunk * 3 * sizeof(double) : 0;
    q.partial = (size_t)blocks * 4 * sizeof(double) * q.nb;
    q.w1 = (size_t)per1 * q.nb;
    q.w2 = (size_t)per2 * q.nb;
    q.w3 = (size_t)per3 * q.nb;
    return SS3D_OK;
}

}  // namespace

extern "C" {

int ss3d_realize_device_noalloc(ss3d_context* c, const ss3d_spectrum* spc, int iseed, int igauss, int64_t r0,
                                int nreal, int nx, int ny, int nz, int prec, void* dev_out, double* dev_stats) {
    SS3D_TRY(check_realize(c, spc, igauss, r0, nreal, nx, ny, nz, prec));
    if (nreal == 0) return SS3D_OK;
    if (!dev_out) return fail(SS3D_EINVAL, "null device output");
    if (!c->scratch) return fail(SS3D_EHIP, "no scratch address on device %d", c->device);
    ss3d_spectrum* sp = const_cast<ss3d_spectrum*>(spc);
    if (!sp->rng[igauss].ready)
        return fail(SS3D_EINVAL,
                    "ss3d_realize_device_noalloc: the spectrum's random-stream tables for igauss=%d are not built; "
                    "one ss3d_realize or ss3d_realize_device call with this igauss builds them",
                    igauss);
    ScratchPlan q;
    SS3D_TRY(plan_scratch(c, sp, nx, ny, nz, prec, nreal, dev_stats == nullptr, q));
    HIP_TRY(hipSetDevice(c->device));
    // windows of the scratch array; buffers this call does not use get none
    Lane& L = c->sl;
    for (DevBuf* b : {&L.wa, &L.wb, &L.pipe[0], &L.pipe[1], &L.partial, &L.partial2, &L.rbase, &L.stats, &L.dout,
                      &L.diag})
        b->set_view(nullptr, 0);
    char* base = static_cast<char*>(c->scratch);
    size_t off = 0;
    auto take = [&](DevBuf& b, size_t bytes) {
        b.set_view(base + off, bytes);
        off += up256(bytes);
    };
    take(L.rbase, q.rbase);
    if (!dev_stats) take(L.stats, q.stats);
    take(L.partial, q.partial);
    if (prec == SS3D_FP32) {
        take(L.pipe[0], q.w1);
        take(L.pipe[1], q.w2);
        take(L.wb, q.w3);
    } else {
        take(L.wa, q.w1);
        take(L.wb, q.w2);
    }
    if (off > kScratchBytes) return fail(SS3D_ETOOBIG, "scratch layout of %zu bytes exceeds %zu", off, kScratchBytes);
    // one stream, no profiling events (event creation is left to the allocating entry points)
    struct Keep {
        ss3d_context* c;
        bool pipeline, profiling;
        ~Keep() {
            c->pipeline = pipeline;
            c->profiling = profiling;
        }
    } keep{c, c->pipeline, c->profiling};
    c->pipeline = false;
    c->profiling = false;
    const int64_t ncrop = (int64_t)nx * ny * nz;
    const size_t esz = prec == SS3D_FP64 ? 8 : 4;
    for (int done = 0; done < nreal; done += q.chunk) {
        const int cnt = std::min(q.chunk, nreal - done);
        char* out = static_cast<char*>(dev_out) + (size_t)done * ncrop * esz;
        double* st = dev_stats ? dev_stats + 3 * (size_t)done : L.stats.as<double>();
        if (prec == SS3D_FP32) {
            SS3D_TRY(enqueue_fp32(c, L, sp, iseed, igauss, r0 + done, cnt, nx, ny, nz, out, st, q.nb));
        } else {
            for (int k = 0; k < cnt; k += q.nb) {
                const int m = std::min(q.nb, cnt - k);
                SS3D_TRY(enqueue_batch(c, L, sp, iseed, igauss, r0 + done + k, m, nx, ny, nz, prec,
                                       out + (size_t)k * ncrop * esz, st + 3 * (size_t)k));
            }
        }
    }
    return SS3D_OK;
}

int ss3d_realize(ss3d_context* c, const ss3d_spectrum* spc, int iseed, int igauss, int64_t r0, int nreal, int nx,
                 int ny, int nz, int prec, void* host_out, double* mean, double* var, double* maximag) {
    SS3D_TRY(check_realize(c, spc, igauss, r0, nreal, nx, ny, nz, prec));
    if (nreal == 0) return SS3D_OK;
    if (!host_out) return fail(SS3D_EINVAL, "null host output");
    HIP_TRY(hipSetDevice(c->device));
    ss3d_spectrum* sp = const_cast<ss3d_spectrum*>(spc);
    const size_t esz = prec == SS3D_FP64 ? 8 : 4;
    const int64_t ncrop = (int64_t)nx * ny * nz;
    // realizations per device round trip: at least one batch, up to ~1 GiB of output
    const int nb = batch_size(sp, nx, ny, nz, prec, nreal);
    const int64_t by_mem = std::max<int64_t>(1, (int64_t(1) << 30) / std::max<int64_t>(ncrop * (int64_t)esz, 1));
    const int ch = (int)std::min<int64_t>(nreal, std::max<int64_t>(nb, by_mem));
    SS3D_TRY(ensure2(c->main, c->main.dout, (size_t)ch * ncrop * esz));
    SS3D_TRY(ensure2(c->main, c->main.stats, (size_t)ch * 3 * sizeof(double)));
    std::vector<double> hs((size_t)ch * 3);
    for (int done = 0; done < nreal; done += ch) {
        const int cnt = std::min(ch, nreal - done);
        SS3D_TRY(ss3d_realize_device(c, sp, iseed, igauss, r0 + done, cnt, nx, ny, nz, prec, c->main.dout.p,
                                     c->main.stats.as<double>()));
        HIP_TRY(hipMemcpyAsync(static_cast<char*>(host_out) + (size_t)done * ncrop * esz, c->main.dout.p,
                               (size_t)cnt * ncrop * esz, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(hs.data(), c->main.stats.p, (size_t)cnt * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (int i = 0; i < cnt; ++i) {
            if (mean) mean[done + i] = hs[3 * i];
            if (var) var[done + i] = hs[3 * i + 1];
            if (maximag) maximag[done + i] = hs[3 * i + 2];
        }
    }
    return SS3D_OK;
}

// Autotuner: for each FP32 pass of this (grid, crop, igauss), time every radix plan
// (largest radix 8, 16, 32) on the actual shape and keep the fastest.  Timing uses the
// per-pass events of the profiler on `reps` batches per plan; results are stored in the
// context and used by every later FP32 call with the same pass lengths.
int ss3d_autotune(ss3d_context* c, const ss3d_spectrum* spc, int nx, int ny, int nz, int igauss, int reps) {
    SS3D_TRY(check_realize(c, spc, igauss, 0, 1, nx, ny, nz, SS3D_FP32));
    HIP_TRY(hipSetDevice(c->device));
    ss3d_spectrum* sp = const_cast<ss3d_spectrum*>(spc);
    const int lx = ilog2(sp->npx), ly = ilog2(sp->npy), lz = ilog2(sp->npz);
    const int Lk[3] = {lz, ly, sp->npx >= 2 ? lx - 1 : 0};
    const int nb = batch_size(sp, nx, ny, nz, SS3D_FP32, 64);
    const int64_t ncrop = (int64_t)nx * ny * nz;
    SS3D_TRY(ensure2(c->main, c->main.dout, (size_t)nb * ncrop * 4));
    SS3D_TRY(ensure2(c->main, c->main.stats, (size_t)nb * 3 * sizeof(double)));
    const bool was = c->profiling;
    // candidates: radix plans 3..5 with the default tile, plans 3..4 with tile variant 1 (where the
    // strided lengths have it); entry = rm + 16 var, as in the tuning table
    const int cand[5] = {3, 4, 5, 3 + 16, 4 + 16};
    std::map<int, double> t[3];
    for (int e : cand) {
        const int rm = e & 15, var = e >> 4;
        bool any = var == 0;
        for (int k = 0; k < 2; ++k) {
            const bool v = var == 1 && fast_tile_variant_exists(Lk[k], rm);
            any = any || v;
            c->tune[{6 + k, Lk[k]}] = rm + (v ? 16 : 0);
        }
        if (!any) continue;
        c->tune[{8, Lk[2]}] = rm;
        SS3D_TRY(enqueue_fp32(c, c->main, sp, 1, igauss, 0, nb, nx, ny, nz, c->main.dout.p, c->main.stats.as<double>()));
        SS3D_TRY(ss3d_context_profile(c, 1));
        for (int r = 0; r < std::max(1, reps); ++r)
            SS3D_TRY(enqueue_fp32(c, c->main, sp, 1, igauss, 0, nb, nx, ny, nz, c->main.dout.p, c->main.stats.as<double>()));
        double ms[SS3D_NPASS_KINDS], by[SS3D_NPASS_KINDS];
        int64_t nl[SS3D_NPASS_KINDS];
        SS3D_TRY(ss3d_context_profile_read(c, ms, by, nl, SS3D_NPASS_KINDS));
        for (int k = 0; k < 3; ++k) {
            const int got = c->tune[{6 + k, Lk[k]}];
            if (nl[6 + k] && (k < 2 || var == 0) && !t[k].count(got)) t[k][got] = ms[6 + k];
        }
    }
    SS3D_TRY(ss3d_context_profile(c, was ? 1 : 0));
    int pick[3] = {5, 5, 5};
    for (int k = 0; k < 3; ++k) {
        double b = 1e30;
        for (auto& kv : t[k])
            if (kv.second < b) {
                b = kv.second;
                pick[k] = kv.first;
            }
    }
    // The fill and the y pass run fused only with the default tiles and plans that have a fused
    // kernel (one plan at one length, plan 3 for a y pass twice as long): where one exists, take
    // the pair with the smallest sum of their times.
    auto fusable = [&](int ez, int ey) { return ez < 16 && ey < 16 && fast_fusable(lz, ez, ly, ey); };
    if (sp->npy > 1 && sp->npz > 1 && !fusable(pick[0], pick[1])) {
        double bestsum = 1e30;
        int pz = 0, py = 0;
        for (auto& a : t[0])
            for (auto& b : t[1])
                if (fusable(a.first, b.first) && a.second + b.second < bestsum) {
                    bestsum = a.second + b.second;
                    pz = a.first;
                    py = b.first;
                }
        if (pz) {
            pick[0] = pz;
            pick[1] = py;
        }
    }
    for (int k = 0; k < 3; ++k) c->tune[{6 + k, Lk[k]}] = pick[k];
    return SS3D_OK;
}

int ss3d_context_set_option(ss3d_context* c, int option, int value) {
    if (!c) return fail(SS3D_EINVAL, "null context");
    if (option == SS3D_OPT_PIPELINE) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipStreamSynchronize(c->main.stream2));
        c->pipeline = value != 0;
        return SS3D_OK;
    }
    if (option == SS3D_OPT_FP64_GENERIC) {
        c->fp64_generic = value != 0;
        return SS3D_OK;
    }
    if (option == SS3D_OPT_FUSE) {
        c->fuse = value != 0;
        return SS3D_OK;
    }
    if (option == SS3D_OPT_BATCH_AUTOTUNE) {
        c->batch_tune = value != 0;
        return SS3D_OK;
    }
    if (option == SS3D_OPT_ZRUN) {
        c->zrun = value != 0;
        return SS3D_OK;
    }
    if (option == SS3D_OPT_SPEC_FP32) {
        c->spec32 = value != 0;
        return SS3D_OK;
    }
    return fail(SS3D_EINVAL, "unknown option %d", option);
}

int ss3d_tuning_get(const ss3d_context* c, int kind, int logn) {
    if (!c) return -1;
    return tuned_rm(c, kind, logn);
}

int ss3d_tuning_set(ss3d_context* c, int kind, int logn, int rm) { return ss3d_tuning_set_tile(c, kind, logn, rm, 0); }

int ss3d_tuning_get_tile(const ss3d_context* c, int kind, int logn, int* rm, int* variant, int* tile, int* lds_pad) {
    if (!c) return fail(SS3D_EINVAL, "null context");
    if (kind < 6 || kind > 8 || logn < 0 || logn > kMaxFastLog)
        return fail(SS3D_EINVAL, "tuning entry out of range (kind 6..8, logn 0..%d)", kMaxFastLog);
    const int r = tuned_rm(c, kind, logn), v = tuned_var(c, kind, logn);
    if (rm) *rm = r;
    if (variant) *variant = v;
    if (tile) *tile = kind == 8 ? fast_c2r_rows(logn, r) : fast_strided_tile(logn, v);
    if (lds_pad) *lds_pad = kind == 8 ? 1 : (fast_strided_pad(logn, v) ? 1 : 0);
    return SS3D_OK;
}

int ss3d_tuning_set_tile(ss3d_context* c, int kind, int logn, int rm, int variant) {
    if (!c) return fail(SS3D_EINVAL, "null context");
    if (kind < 6 || kind > 8 || logn < 0 || logn > kMaxFastLog || rm < 3 || rm > 5)
        return fail(SS3D_EINVAL, "tuning entry out of range (kind 6..8, logn 0..%d, rm 3..5)", kMaxFastLog);
    if (variant != 0 && (variant != 1 || kind == 8 || !fast_tile_variant_exists(logn, rm)))
        return fail(SS3D_EINVAL, "no tile variant %d for pass kind %d, length 2^%d, plan %d", variant, kind, logn, rm);
    c->tune[{kind, logn}] = rm + 16 * variant;
    return SS3D_OK;
}

}  // extern "C"

#include "ss3d_server.inc"
#include "ss3d_batch.inc"
#include "ss3d_vario.inc"
#include "ss3d_output.inc"
#include "ss3d_specfile.inc"
#include "ss3d_cond.inc"
