// lekf_capi.hip -- C ABI of the LETKF library: per-point drop-in, batched
// solve, and the das_letkf-level handle that keeps the ensemble resident on
// the device.  Host code only orchestrates: every array operation of the
// analysis runs in the kernels of lekf_solve.hip / lekf_cull.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "lekf_cull.hpp"
#include "lekf_radar.hpp"
#include "lekf_dev.hpp"
#include "lekf_regional.hpp"
#include "lekf_obsope.hpp"
#include "lekf_relax.hpp"
#include "lekf_l96.hpp"
#include "lekf_tune.hpp"
#include "lekf_ktime.hpp"

#define LEKF_VERSION "lekf-mi355x 0.1 (gfx950)"

LekfTune g_lekf_tune;

namespace {

thread_local char g_errmsg[256];
// launch tuning; the epoch counts changes so captured graphs can be retaken
std::atomic<long> g_tune_epoch{0};

void set_err(hipError_t e, const char* what)
{
    snprintf(g_errmsg, sizeof(g_errmsg), "%s: %s", what, hipGetErrorString(e));
}

#define HIPCHK(x)                                    \
    do {                                             \
        hipError_t e_ = (x);                         \
        if (e_ != hipSuccess) {                      \
            set_err(e_, #x);                         \
            return LEKF_ERR_HIP;                     \
        }                                            \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes == 0) bytes = 8;
        if (bytes <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <class T>
    T* as() const
    {
        return (T*)p;
    }
};

__global__ void iota_kernel(int* a, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// reference column-major (nij, nlev, ne, nv) <-> device rows [(lev*nij+ij)*nv + v][m]
__global__ void ens3_to_dev(const double* ref, double* dev, int nij, int nlev, int ne, int nv)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tot = (long)nij * nlev * ne * nv;
    if (t >= tot) return;
    // t enumerates the reference layout (coalesced reads)
    const int ij = (int)(t % nij);
    long r = t / nij;
    const int lev = (int)(r % nlev);
    r /= nlev;
    const int m = (int)(r % ne);
    const int v = (int)(r / ne);
    dev[(((long)lev * nij + ij) * nv + v) * ne + m] = ref[t];
}

__global__ void ens3_from_dev(const double* dev, double* ref, int nij, int nlev, int ne, int nv)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tot = (long)nij * nlev * ne * nv;
    if (t >= tot) return;
    const int ij = (int)(t % nij);
    long r = t / nij;
    const int lev = (int)(r % nlev);
    r /= nlev;
    const int m = (int)(r % ne);
    const int v = (int)(r / ne);
    ref[t] = dev[(((long)lev * nij + ij) * nv + v) * ne + m];
}

// Lorenz-96 obs set-up: mean of H(x), perturbations, departure, R = err^2
// (lorenz96/letkf/letkf.f90:185-214)
__global__ void l96_obs_kernel(int ny, int ne, const double* hxf, const double* y,
                               const double* err, double* hdxf, float* hdxf32, double* dep,
                               double* rdiag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ny) return;
    double m = hxf[i];
    for (int j = 1; j < ne; ++j) m += hxf[i + (long)j * ny];
    m /= (double)ne;
    for (int j = 0; j < ne; ++j) {
        const double v = hxf[i + (long)j * ny] - m;
        hdxf[(long)i * ne + j] = v;
        if (hdxf32) hdxf32[(long)i * ne + j] = (float)v;
    }
    dep[i] = y[i] - m;
    rdiag[i] = err[i] * err[i];
}

int mode_of_prec(int prec) { return prec == LEKF_PREC_F32 ? 1 : (prec == LEKF_PREC_BF16X3 ? 2 : 0); }

// ---------------------------------------------------------------------------
// per-thread context of the per-point drop-in and of lekf_core_batch
// ---------------------------------------------------------------------------
struct CoreCtx {
    bool init = false;
    hipStream_t st = nullptr;
    DevBuf h, rd, rl, dp, idx, off, infl, tr, tm, pa, stat, job, gws;
    int iota_n = 0;
    std::map<hipStream_t, DevBuf*> jobs_by_stream, ws_by_stream;
};
thread_local CoreCtx t_core;

int core_init()
{
    if (t_core.init) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return LEKF_ERR_NODEV;
    HIPCHK(hipStreamCreateWithFlags(&t_core.st, hipStreamNonBlocking));
    t_core.init = true;
    return 0;
}

// ---------------------------------------------------------------------------
// Device scratch of lekf_core_batch_noalloc: the jobs' descriptors, then the
// Gram workspace, in a file-scope array of fixed size (each GPU holds its own
// copy for the life of the process).  Host code never uses the array's name
// as an address: the device address comes from hipGetSymbolAddress, once per
// device, with that device current.  No solve kernel reads workspace it has
// not written in the same call, so nothing is cleared between calls; calls
// that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[LEKF_NOALLOC_BYTES];

// descriptors first, the workspace from a 4 KiB boundary after them
constexpr size_t kNoallocJobBytes =
    ((size_t)LEKF_NOALLOC_MAX_JOBS * sizeof(lekf_job) + 4095) & ~(size_t)4095;
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// device address of this device's scratch (nullptr and an error message on failure)
void* scratch_ptr()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        snprintf(g_errmsg, sizeof(g_errmsg), "lekf_core_batch_noalloc: no usable device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            snprintf(g_errmsg, sizeof(g_errmsg), "hipGetSymbolAddress(scratch): %s",
                     hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

// One solve launch; ws is grown to the Gram workspace the class needs.
int run_solve(const lekf_job* djobs, int njobs, long npt, int kmax, int nvmax, int mode,
              DevBuf& ws, hipStream_t st)
{
    const size_t need = lekf_solve_ws_bytes(npt, kmax);
    if (need) HIPCHK(ws.ensure(need));
    HIPCHK(lekf_solve_launch(djobs, njobs, npt, kmax, nvmax, mode, need ? ws.p : nullptr,
                             need ? ws.n : 0, st));
    return 0;
}

// state rows a job's points apply their transform to (0 without the apply)
int job_rows(const lekf_job& j) { return (j.flags & LEKF_F_APPLY) ? j.nv3d + j.nv2d : 0; }

// The grouped launches of launch_grouped over caller-provided device memory:
// descriptors at dj (room for J.size()), workspace ws of ws_bytes.  Nothing is
// allocated.  Returns LEKF_ERR_SIZE when a class's point record does not fit.
int launch_grouped_fixed(std::vector<lekf_job> J, int mode, lekf_job* dj, void* ws,
                         size_t ws_bytes, hipStream_t st)
{
    std::stable_sort(J.begin(), J.end(), [](const lekf_job& a, const lekf_job& b) {
        return lekf_solve_class(a.k) < lekf_solve_class(b.k);
    });
    std::vector<size_t> starts;
    for (size_t i = 0; i < J.size(); ++i)
        if (i == 0 || lekf_solve_class(J[i].k) != lekf_solve_class(J[i - 1].k)) starts.push_back(i);
    starts.push_back(J.size());
    std::vector<long> tot(starts.size(), 0);
    std::vector<int> kmax(starts.size(), 0), nvmax(starts.size(), 0);
    for (size_t g = 0; g + 1 < starts.size(); ++g)
        for (size_t i = starts[g]; i < starts[g + 1]; ++i) {
            J[i].pt_begin = tot[g];
            tot[g] += J[i].npt;
            kmax[g] = std::max(kmax[g], J[i].k);
            nvmax[g] = std::max(nvmax[g], job_rows(J[i]));
        }
    for (size_t g = 0; g + 1 < starts.size(); ++g)
        if (tot[g] > 0 && lekf_solve_ws_bytes(1, kmax[g]) > ws_bytes) {
            snprintf(g_errmsg, sizeof(g_errmsg),
                     "lekf_core_batch_noalloc: a point record of k = %d needs %zu bytes, the "
                     "workspace holds %zu", kmax[g], lekf_solve_ws_bytes(1, kmax[g]), ws_bytes);
            return LEKF_ERR_SIZE;
        }
    HIPCHK(hipMemcpyAsync(dj, J.data(), sizeof(lekf_job) * J.size(), hipMemcpyHostToDevice, st));
    for (size_t g = 0; g + 1 < starts.size(); ++g)
        HIPCHK(lekf_solve_launch(dj + starts[g], (int)(starts[g + 1] - starts[g]), tot[g], kmax[g],
                                 nvmax[g], mode, kmax[g] > 32 ? ws : nullptr,
                                 kmax[g] > 32 ? ws_bytes : 0, st));
    return 0;
}

// Launch the solves of jobs (all of one arithmetic mode) grouped by kernel
// class: descriptors are uploaded once, sorted by class, and each class is
// one launch whose points are numbered from 0.
int launch_grouped(std::vector<lekf_job> J, int mode, DevBuf& db, DevBuf& ws, hipStream_t st)
{
    std::stable_sort(J.begin(), J.end(), [](const lekf_job& a, const lekf_job& b) {
        return lekf_solve_class(a.k) < lekf_solve_class(b.k);
    });
    HIPCHK(db.ensure(sizeof(lekf_job) * J.size()));
    std::vector<size_t> starts;
    for (size_t i = 0; i < J.size(); ++i)
        if (i == 0 || lekf_solve_class(J[i].k) != lekf_solve_class(J[i - 1].k)) starts.push_back(i);
    starts.push_back(J.size());
    std::vector<long> tot(starts.size(), 0);
    std::vector<int> kmax(starts.size(), 0), nvmax(starts.size(), 0);
    for (size_t g = 0; g + 1 < starts.size(); ++g)
        for (size_t i = starts[g]; i < starts[g + 1]; ++i) {
            J[i].pt_begin = tot[g];
            tot[g] += J[i].npt;
            kmax[g] = std::max(kmax[g], J[i].k);
            nvmax[g] = std::max(nvmax[g], job_rows(J[i]));
        }
    HIPCHK(hipMemcpyAsync(db.p, J.data(), sizeof(lekf_job) * J.size(), hipMemcpyHostToDevice, st));
    for (size_t g = 0; g + 1 < starts.size(); ++g) {
        const int rc = run_solve(db.as<lekf_job>() + starts[g], (int)(starts[g + 1] - starts[g]),
                                 tot[g], kmax[g], nvmax[g], mode, ws, st);
        if (rc) return rc;
    }
    return 0;
}

int ensure_iota(DevBuf& b, int& have, int n, hipStream_t st)
{
    if (n <= have) return 0;
    HIPCHK(b.ensure(sizeof(int) * (size_t)n));
    hipLaunchKernelGGL(iota_kernel, dim3((n + 255) / 256), dim3(256), 0, st, b.as<int>(), n);
    HIPCHK(hipGetLastError());
    have = n;
    return 0;
}

}  // namespace

// ===========================================================================
extern "C" {

const char* lekf_version(void) { return LEKF_VERSION; }

// --- device memory, streams and events for hosts without a GPU runtime of their own
void* lekf_dev_malloc(size_t bytes)
{
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return nullptr;
    return p;
}

int lekf_dev_free(void* p)
{
    HIPCHK(hipFree(p));
    return LEKF_OK;
}

// streaming device copy in 16-byte words (one read and one write of each byte)
__global__ void copy16_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

int lekf_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream)
{
    if (kind == 2 && bytes % 16 == 0 && ((uintptr_t)dst | (uintptr_t)src) % 16 == 0) {
        const size_t n = bytes / 16;
        if (n) {
            const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
            hipLaunchKernelGGL(copy16_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream,
                               (const uint4*)src, (uint4*)dst, n);
            HIPCHK(hipGetLastError());
        }
        return LEKF_OK;
    }
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                          : kind == 1 ? hipMemcpyDeviceToHost
                                      : hipMemcpyDeviceToDevice;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream));
    if (kind != 2) HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    return LEKF_OK;
}

int lekf_memset(void* dst, int value, size_t bytes, void* stream)
{
    HIPCHK(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
    return LEKF_OK;
}

void* lekf_stream_create(void)
{
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
    return (void*)s;
}

int lekf_stream_destroy(void* stream)
{
    HIPCHK(hipStreamDestroy((hipStream_t)stream));
    return LEKF_OK;
}

int lekf_stream_sync(void* stream)
{
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    return LEKF_OK;
}

void* lekf_event_create(void)
{
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return (void*)e;
}

int lekf_event_destroy(void* ev)
{
    HIPCHK(hipEventDestroy((hipEvent_t)ev));
    return LEKF_OK;
}

int lekf_event_record(void* ev, void* stream)
{
    HIPCHK(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return LEKF_OK;
}

int lekf_event_elapsed_ms(void* ev0, void* ev1, float* ms)
{
    HIPCHK(hipEventSynchronize((hipEvent_t)ev1));
    HIPCHK(hipEventElapsedTime(ms, (hipEvent_t)ev0, (hipEvent_t)ev1));
    return LEKF_OK;
}

int lekf_event_query(void* ev)
{
    const hipError_t e = hipEventQuery((hipEvent_t)ev);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    set_err(e, "hipEventQuery");
    return -LEKF_ERR_HIP;
}

int lekf_event_sync(void* ev)
{
    HIPCHK(hipEventSynchronize((hipEvent_t)ev));
    return LEKF_OK;
}

const char* lekf_last_error(void) { return g_errmsg; }

int lekf_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int lekf_tune_set(int param, long value)
{
    switch (param) {
    case LEKF_TUNE_GRAM_WAVES:
        if (value != 2 && value != 4 && value != 8) return LEKF_ERR_ARG;
        g_lekf_tune.gram_waves = (int)value;
        break;
    case LEKF_TUNE_GRAM_WS_MB:
        if (value < 16 || value > 65536) return LEKF_ERR_ARG;
        g_lekf_tune.gram_ws_bytes = (size_t)value << 20;
        break;
    case LEKF_TUNE_SEL_PER_CU:
        if (value < 1 || value > 4) return LEKF_ERR_ARG;
        g_lekf_tune.sel_per_cu = (int)value;
        break;
    case LEKF_TUNE_EIG_SPLIT:
        if (value != 0 && value != 1) return LEKF_ERR_ARG;
        g_lekf_tune.eig_split = (int)value;
        break;
    case LEKF_TUNE_SPLIT_RCAP:
        if (value < 0) return LEKF_ERR_ARG;
        g_lekf_tune.split_rcap = value;
        break;
    case LEKF_TUNE_MIXED_KAPPA:
        if (value < 0 || value > 1000000000000L) return LEKF_ERR_ARG;
        g_lekf_tune.mixed_kappa = value;
        break;
    default: return LEKF_ERR_ARG;
    }
    g_tune_epoch.fetch_add(1);
    return LEKF_OK;
}

}  // extern "C"

// --- per-kernel timing (measurement) ---------------------------------------
namespace {
std::atomic<int> g_ktime_on{0};
std::mutex g_ktime_mu;
struct KRec {
    const char* name;
    hipStream_t st;
    hipEvent_t a, b;
};
std::vector<KRec> g_ktime_recs;       // event pairs in launch order
std::vector<hipEvent_t> g_ktime_pool; // events of cleared pairs, for reuse

void ktime_clear_locked()
{
    for (KRec& r : g_ktime_recs) {
        if (r.a) g_ktime_pool.push_back(r.a);
        if (r.b) g_ktime_pool.push_back(r.b);
    }
    g_ktime_recs.clear();
}
}  // namespace

void lekf_ktime_mark(const char* name, hipStream_t st, bool begin)
{
    if (!g_ktime_on.load(std::memory_order_relaxed)) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    std::lock_guard<std::mutex> lk(g_ktime_mu);
    hipEvent_t e = nullptr;
    if (!g_ktime_pool.empty()) {
        e = g_ktime_pool.back();
        g_ktime_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        return;
    }
    if (hipEventRecord(e, st) != hipSuccess) {
        g_ktime_pool.push_back(e);
        return;
    }
    if (begin) {
        g_ktime_recs.push_back({name, st, e, nullptr});
        return;
    }
    for (size_t i = g_ktime_recs.size(); i-- > 0;) {
        KRec& r = g_ktime_recs[i];
        if (!r.b && r.st == st && std::strcmp(r.name, name) == 0) {
            r.b = e;
            return;
        }
    }
    g_ktime_pool.push_back(e);
}

extern "C" {

long lekf_tune_get(int param)
{
    switch (param) {
    case LEKF_TUNE_GRAM_WAVES: return g_lekf_tune.gram_waves;
    case LEKF_TUNE_GRAM_WS_MB: return (long)(g_lekf_tune.gram_ws_bytes >> 20);
    case LEKF_TUNE_SEL_PER_CU: return g_lekf_tune.sel_per_cu;
    case LEKF_TUNE_EIG_SPLIT: return g_lekf_tune.eig_split;
    case LEKF_TUNE_SPLIT_RCAP: return g_lekf_tune.split_rcap;
    case LEKF_TUNE_MIXED_KAPPA: return g_lekf_tune.mixed_kappa;
    default: return -LEKF_ERR_ARG;
    }
}

int lekf_ktime_enable(int on)
{
    std::lock_guard<std::mutex> lk(g_ktime_mu);
    if (on) ktime_clear_locked();
    g_ktime_on.store(on ? 1 : 0);
    return LEKF_OK;
}

int lekf_ktime_read(const char* name, double* total_ms, long* launches)
{
    if (!name || !total_ms || !launches) return LEKF_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_ktime_mu);
    double t = 0.0;
    long n = 0;
    for (const KRec& r : g_ktime_recs) {
        if (!r.b || std::strcmp(r.name, name) != 0) continue;
        HIPCHK(hipEventSynchronize(r.b));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
        t += ms;
        ++n;
    }
    *total_ms = t;
    *launches = n;
    return LEKF_OK;
}

int lekf_mem_info(size_t* free_bytes, size_t* total_bytes)
{
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return LEKF_OK;
}

long lekf_solve_workspace_bytes(long npt, int k)
{
    if (npt <= 0 || k < 2 || k > LEKF_KMAX) return 0;
    return (long)lekf_solve_ws_bytes(npt, k);
}

int lekf_core_f64(int ne, int nobs, int nobsl, const double* hdxb, const double* rdiag,
                  const double* rloc, const double* dep, double* parm_infl, double* trans,
                  double* transm, double* pao, int rdiag_wloc, const double* minfl)
{
    if (ne < 2 || ne > LEKF_KMAX || nobsl < 0 || nobs < nobsl || !parm_infl || !trans)
        return LEKF_ERR_SIZE;
    if (nobsl > 0 && (!hdxb || !rdiag || !rloc || !dep)) return LEKF_ERR_ARG;
    int rc = core_init();
    if (rc) return rc;
    CoreCtx& c = t_core;
    hipStream_t st = c.st;
    const size_t p = (size_t)nobsl, k = (size_t)ne;
    HIPCHK(c.h.ensure(sizeof(double) * (p ? p : 1) * k));
    HIPCHK(c.rd.ensure(sizeof(double) * (p ? p : 1)));
    HIPCHK(c.rl.ensure(sizeof(double) * (p ? p : 1)));
    HIPCHK(c.dp.ensure(sizeof(double) * (p ? p : 1)));
    HIPCHK(c.off.ensure(sizeof(long) * 2));
    HIPCHK(c.infl.ensure(sizeof(double) * 2));
    HIPCHK(c.tr.ensure(sizeof(double) * k * k));
    HIPCHK(c.tm.ensure(sizeof(double) * k));
    HIPCHK(c.pa.ensure(sizeof(double) * k * k));
    HIPCHK(c.stat.ensure(sizeof(int)));
    HIPCHK(c.job.ensure(sizeof(lekf_job)));
    rc = ensure_iota(c.idx, c.iota_n, (int)(p ? p : 1), st);
    if (rc) return rc;
    if (p) {
        HIPCHK(hipMemcpy2DAsync(c.h.p, p * sizeof(double), hdxb, (size_t)nobs * sizeof(double),
                                p * sizeof(double), k, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c.rd.p, rdiag, p * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c.rl.p, rloc, p * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c.dp.p, dep, p * sizeof(double), hipMemcpyHostToDevice, st));
    }
    const long offs[2] = {0, (long)p};
    HIPCHK(hipMemcpyAsync(c.off.p, offs, sizeof(offs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(c.infl.p, parm_infl, sizeof(double), hipMemcpyHostToDevice, st));
    lekf_job J;
    memset(&J, 0, sizeof(J));
    J.pt_begin = 0;
    J.npt = 1;
    J.k = ne;
    J.flags = (rdiag_wloc ? LEKF_F_WLOC : 0) | (minfl ? LEKF_F_MINFL : 0) |
              (transm ? LEKF_F_TRANSM : 0);
    J.minfl = minfl ? *minfl : 0.0;
    J.hdxf = c.h.as<double>();
    J.y_so = 1;
    J.y_sm = (long)(p ? p : 1);
    J.rdiag = c.rd.as<double>();
    J.dep = c.dp.as<double>();
    J.loc_off = c.off.as<long>();
    J.loc_idx = c.idx.as<int>();
    J.loc_rho = c.rl.as<double>();
    J.infl_in = c.infl.as<double>();
    J.infl_out = c.infl.as<double>() + 1;
    J.trans = c.tr.as<double>();
    J.transm = transm ? c.tm.as<double>() : nullptr;
    J.pao = pao ? c.pa.as<double>() : nullptr;
    J.status = c.stat.as<int>();
    HIPCHK(hipMemcpyAsync(c.job.p, &J, sizeof(J), hipMemcpyHostToDevice, st));
    rc = run_solve(c.job.as<lekf_job>(), 1, 1, ne, 0, 0, c.gws, st);
    if (rc) return rc;
    int pst = 0;
    HIPCHK(hipMemcpyAsync(trans, c.tr.p, sizeof(double) * k * k, hipMemcpyDeviceToHost, st));
    if (transm) HIPCHK(hipMemcpyAsync(transm, c.tm.p, sizeof(double) * k, hipMemcpyDeviceToHost, st));
    if (pao) HIPCHK(hipMemcpyAsync(pao, c.pa.p, sizeof(double) * k * k, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(parm_infl, c.infl.as<double>() + 1, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&pst, c.stat.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (pst & LEKF_PT_NOCONV) return LEKF_ERR_EIGEN;
    if (pst & LEKF_PT_RANKDEF) return LEKF_WARN_RANKDEF;
    return LEKF_OK;
}

int lekf_core_batch(const lekf_job* jobs, int njobs, int prec, void* stream)
{
    if (njobs <= 0 || !jobs) return LEKF_ERR_ARG;
    int rc = core_init();
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    std::vector<lekf_job> J(jobs, jobs + njobs);
    long tot = 0;
    int kmax = 0;
    const int mode = mode_of_prec(prec);
    for (auto& j : J) {
        if (j.k < 2 || j.k > LEKF_KMAX || j.npt < 0) return LEKF_ERR_SIZE;
        if ((j.flags & LEKF_F_APPLY) && (j.nv3d + j.nv2d > LEKF_NVMAX || j.nij <= 0))
            return LEKF_ERR_SIZE;
        if (mode != 0 && !j.hdxf32) return LEKF_ERR_ARG;
        if (!j.loc_off || (j.npt > 0 && (!j.loc_idx || !j.loc_rho))) return LEKF_ERR_ARG;
        j.pt_begin = tot;
        tot += j.npt;
        kmax = std::max(kmax, j.k);
    }
    if (tot == 0) return LEKF_OK;
    DevBuf*& db = t_core.jobs_by_stream[st];
    if (!db) db = new DevBuf();
    DevBuf*& wb = t_core.ws_by_stream[st];
    if (!wb) wb = new DevBuf();
    return launch_grouped(J, mode, *db, *wb, st);
}

int lekf_core_batch_noalloc(const lekf_job* jobs, int njobs, int prec, void* stream)
{
    if (njobs <= 0 || !jobs) return LEKF_ERR_ARG;
    if (njobs > LEKF_NOALLOC_MAX_JOBS) {
        snprintf(g_errmsg, sizeof(g_errmsg), "lekf_core_batch_noalloc: %d jobs, at most %d", njobs,
                 LEKF_NOALLOC_MAX_JOBS);
        return LEKF_ERR_SIZE;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return LEKF_ERR_NODEV;
    std::vector<lekf_job> J(jobs, jobs + njobs);
    long tot = 0;
    const int mode = mode_of_prec(prec);
    for (auto& j : J) {
        if (j.k < 2 || j.k > LEKF_KMAX || j.npt < 0) return LEKF_ERR_SIZE;
        if ((j.flags & LEKF_F_APPLY) && (j.nv3d + j.nv2d > LEKF_NVMAX || j.nij <= 0))
            return LEKF_ERR_SIZE;
        if (mode != 0 && !j.hdxf32) return LEKF_ERR_ARG;
        if (!j.loc_off || (j.npt > 0 && (!j.loc_idx || !j.loc_rho))) return LEKF_ERR_ARG;
        tot += j.npt;
    }
    if (tot == 0) return LEKF_OK;
    unsigned char* base = (unsigned char*)scratch_ptr();
    if (!base) return LEKF_ERR_HIP;
    return launch_grouped_fixed(J, mode, (lekf_job*)base, base + kNoallocJobBytes,
                                (size_t)LEKF_NOALLOC_BYTES - kNoallocJobBytes, (hipStream_t)stream);
}

int lekf_obs_local_l96(int nx, int ny, const int* gidx, double xlocal, long* loc_off, int* loc_cnt,
                       int* loc_idx, double* loc_rho, void* stream)
{
    if (nx <= 0 || ny < 0) return LEKF_ERR_SIZE;
    L96CullArgs A;
    A.nx = nx;
    A.ny = ny;
    A.gidx = gidx;
    A.xlocal = xlocal;
    A.xmaxloc = xlocal * 2.0 * std::sqrt(10.0 / 3.0);
    A.loc_off = loc_off;
    A.loc_cnt = loc_cnt;
    A.loc_idx = loc_idx;
    A.loc_rho = loc_rho;
    HIPCHK(lekf_cull_l96_launch(A, (hipStream_t)stream));
    return LEKF_OK;
}

}  // extern "C"

// ===========================================================================
// das_letkf-level handle
// ===========================================================================
struct lekf_handle {
    int kind = 0;  // 0 global, 1 Lorenz-96, 2 regional
    hipStream_t st = nullptr;
    bool own_stream = true;
    int ne = 0, nij = 0, nlev = 1, nv3d = 1, nv2d = 0;
    int prec = 0;
    int infl_mode = 0;
    double infl_value = 1.0;
    lekf_global_cfg g{};
    lekf_l96_cfg l{};
    std::vector<double> lon, lat, sig;
    double dist_zero = 0, dist_zerov = 0, dlat_zero = 0;
    // device grid
    DevBuf d_lon, d_lat, d_sig, d_col_lon, d_col_lat, d_col_dlz, d_var_local, d_logpfm;
    // state
    DevBuf x3b, x3a, x2b, x2a, upd, infl_in, infl_out, infl2, stage;
    bool have_ens = false;
    // obs
    int nobs = 0, nbinned = 0;
    DevBuf r_elm, r_lon, r_lat, r_lev, r_dat, r_err, r_qc, r_hxf, r_y, r_gidx;
    DevBuf tmp_mean, tmp_dep, keys, cell_cnt, perm, o_hdxf, o_hdxf32, o_dep, o_rdiag, o_err,
        o_typ, o_sin, o_cos, o_lonr, o_lnlev, o_lndat, nobsgrd;
    bool have_obs = false;
    // replay of the cycle's launches as one HIP graph (lekf_set_graph)
    bool use_graph = false, graph_ok = false;
    hipGraphExec_t gexec = nullptr;
    long graph_epoch = 0;  // tuning epoch the graph was captured under
    // relaxation pass after the solves (lekf_set_relax)
    int relax_mode = LEKF_RELAX_NONE, relax_iq = -1;
    double relax_alpha = 0.0, relax_qmax = 0.0, relax_width = 0.0;
    // local lists
    DevBuf cand_cnt, cand_off, loc_off, loc_cnt, loc_idx, loc_rho, status;
    long loc_cap = 0;
    // job
    DevBuf djob;
    lekf_job job{};
    long bytes = 0;
    // regional grid (kind 2)
    lekf_regional_cfg r{};
    std::vector<double> zlev;
    int nbkt = 1;
    double dlon_zero = 0;  // regional rectangle half-width in grid units (dlat_zero above)
    DevBuf d_zlev, cellcum, cap_n, cap_off, o_ri, o_rj, o_z, o_bkt, r_dep, r_bkt, r_z;
    DevBuf d_hgt;                    // level heights (lekf_set_heights), radar operator
    DevBuf gws;                      // Gram workspace of the k > 32 solve
    DevBuf sel_key, sel_ob, sel_rho;  // per-slot bucket streams of the capped regional selection
    DevBuf sel_gkey, sel_gpos, sel_gslot;  // per-slot kept sets of caps above LEKF_CAP_LDS
    // the sorted obs bucket by bucket, for the LDS-resident capped selection
    DevBuf b_cnt, b_cc, b_rec, b_o, b_err2min;
    DevBuf wc_so;                    // its slot -> obs maps, one per resident wave
    DevBuf maxcand;
    long slot_cap = 0;
    int sel_slots = 0;
    int gcap = 0;  // global grid: the cap in effect (nobs_max, at most the longest raw list)
    // global grid with nobs_max: raw lists per column chunk, capped lists at b * nobs_max
    DevBuf capl_off, capl_cnt, capl_idx, capl_rho;
    std::vector<int> gchunk_cols;    // column chunk boundaries
    std::vector<long> gchunk_base;   // cand_off * nlev at the boundaries
    std::vector<long> chunk_pts;  // chunk boundaries (points)
    std::vector<long> chunk_off;  // list offsets at the boundaries
    long list_budget = 512L * 1024 * 1024;
    int held_chunk = -1;  // chunk whose lists the list buffer holds
};

namespace {

long handle_bytes(const lekf_handle* h)
{
    const DevBuf* bs[] = {&h->d_lon, &h->d_lat, &h->d_sig, &h->d_col_lon, &h->d_col_lat,
                          &h->d_col_dlz, &h->d_var_local, &h->d_logpfm, &h->x3b, &h->x3a,
                          &h->x2b, &h->x2a, &h->upd, &h->infl_in, &h->infl_out, &h->infl2,
                          &h->stage, &h->r_elm, &h->r_lon, &h->r_lat, &h->r_lev, &h->r_dat,
                          &h->r_err, &h->r_qc, &h->r_hxf, &h->r_y, &h->r_gidx, &h->tmp_mean,
                          &h->tmp_dep, &h->keys, &h->cell_cnt, &h->perm, &h->o_hdxf,
                          &h->o_hdxf32, &h->o_dep, &h->o_rdiag, &h->o_err, &h->o_typ, &h->o_sin,
                          &h->o_cos, &h->o_lonr, &h->o_lnlev, &h->o_lndat, &h->nobsgrd,
                          &h->cand_cnt, &h->cand_off, &h->loc_off, &h->loc_cnt, &h->loc_idx,
                          &h->loc_rho, &h->status, &h->djob, &h->d_zlev, &h->cellcum,
                          &h->cap_n, &h->cap_off, &h->o_ri, &h->o_rj, &h->o_z, &h->o_bkt,
                          &h->r_dep, &h->r_bkt, &h->r_z, &h->gws, &h->sel_key,
                          &h->sel_ob, &h->sel_rho, &h->sel_gkey, &h->sel_gpos, &h->sel_gslot,
                          &h->maxcand, &h->capl_off, &h->capl_cnt,
                          &h->capl_idx, &h->capl_rho, &h->d_hgt, &h->b_cnt, &h->b_cc,
                          &h->b_rec, &h->b_o, &h->b_err2min, &h->wc_so};
    long s = 0;
    for (auto* b : bs) s += (long)b->n;
    return s;
}

void free_all(lekf_handle* h)
{
    DevBuf* bs[] = {&h->d_lon, &h->d_lat, &h->d_sig, &h->d_col_lon, &h->d_col_lat, &h->d_col_dlz,
                    &h->d_var_local, &h->d_logpfm, &h->x3b, &h->x3a, &h->x2b, &h->x2a, &h->upd,
                    &h->infl_in, &h->infl_out, &h->infl2, &h->stage, &h->r_elm, &h->r_lon,
                    &h->r_lat, &h->r_lev, &h->r_dat, &h->r_err, &h->r_qc, &h->r_hxf, &h->r_y,
                    &h->r_gidx, &h->tmp_mean, &h->tmp_dep, &h->keys, &h->cell_cnt, &h->perm,
                    &h->o_hdxf, &h->o_hdxf32, &h->o_dep, &h->o_rdiag, &h->o_err, &h->o_typ,
                    &h->o_sin, &h->o_cos, &h->o_lonr, &h->o_lnlev, &h->o_lndat, &h->nobsgrd,
                    &h->cand_cnt, &h->cand_off, &h->loc_off, &h->loc_cnt, &h->loc_idx, &h->loc_rho,
                    &h->status, &h->djob, &h->d_zlev, &h->cellcum, &h->cap_n, &h->cap_off,
                    &h->o_ri, &h->o_rj, &h->o_z, &h->o_bkt, &h->r_dep, &h->r_bkt, &h->r_z,
                    &h->gws, &h->sel_key, &h->sel_ob, &h->sel_rho, &h->sel_gkey, &h->sel_gpos,
                    &h->sel_gslot, &h->maxcand, &h->capl_off,
                    &h->capl_cnt, &h->capl_idx, &h->capl_rho, &h->d_hgt, &h->b_cnt, &h->b_cc,
                    &h->b_rec, &h->b_o, &h->b_err2min, &h->wc_so};
    for (auto* b : bs) b->release();
}

// kept sets of a cap above LEKF_CAP_LDS: 2 * cap entries per selection slot
hipError_t ensure_cap_sets(lekf_handle* h, int cap)
{
    if (cap <= LEKF_CAP_LDS) return hipSuccess;
    const size_t n = (size_t)h->sel_slots * 2 * (size_t)cap;
    hipError_t e;
    if ((e = h->sel_gkey.ensure(sizeof(unsigned long long) * n)) != hipSuccess) return e;
    if ((e = h->sel_gpos.ensure(sizeof(int) * n)) != hipSuccess) return e;
    return h->sel_gslot.ensure(sizeof(int) * n);
}

int upload(DevBuf& b, const void* src, size_t bytes, hipStream_t st)
{
    HIPCHK(b.ensure(bytes));
    if (bytes && src) HIPCHK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st));
    return 0;
}

int alloc_state(lekf_handle* h)
{
    const size_t n3 = (size_t)h->nij * h->nlev * h->nv3d * h->ne;
    const size_t n2 = (size_t)h->nij * h->nv2d * h->ne;
    HIPCHK(h->x3b.ensure(sizeof(double) * n3));
    HIPCHK(h->x3a.ensure(sizeof(double) * n3));
    HIPCHK(h->x2b.ensure(sizeof(double) * (n2 ? n2 : 1)));
    HIPCHK(h->x2a.ensure(sizeof(double) * (n2 ? n2 : 1)));
    const size_t npt = (size_t)h->nij * h->nlev;
    HIPCHK(h->infl_in.ensure(sizeof(double) * npt));
    HIPCHK(h->infl_out.ensure(sizeof(double) * npt));
    HIPCHK(h->infl2.ensure(sizeof(double) * (size_t)h->nij * (h->nv2d ? h->nv2d : 1)));
    HIPCHK(h->status.ensure(sizeof(int) * npt));
    HIPCHK(h->loc_off.ensure(sizeof(long) * (npt + 1)));
    HIPCHK(h->loc_cnt.ensure(sizeof(int) * npt));
    // initial inflation field
    std::vector<double> f(npt, h->infl_value);
    HIPCHK(hipMemcpyAsync(h->infl_in.p, f.data(), sizeof(double) * npt, hipMemcpyHostToDevice,
                          h->st));
    std::vector<double> f2((size_t)h->nij * (h->nv2d ? h->nv2d : 1), h->infl_value);
    HIPCHK(hipMemcpyAsync(h->infl2.p, f2.data(), sizeof(double) * f2.size(), hipMemcpyHostToDevice,
                          h->st));
    // update masks (SPEEDY q rule, speedy/letkf/letkf_tools.f90:172-177)
    std::vector<uint32_t> m(h->nlev, 0xffffffffu);
    if (h->kind == 0 && h->g.iq >= 0 && h->g.q_top_lev > 0)
        for (int l = h->g.q_top_lev - 1; l < h->nlev; ++l) m[l] &= ~(1u << h->g.iq);
    HIPCHK(h->upd.ensure(sizeof(uint32_t) * h->nlev));
    HIPCHK(hipMemcpyAsync(h->upd.p, m.data(), sizeof(uint32_t) * h->nlev, hipMemcpyHostToDevice,
                          h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

void fill_job(lekf_handle* h)
{
    lekf_job& J = h->job;
    memset(&J, 0, sizeof(J));
    J.npt = h->nij * h->nlev;
    J.k = h->ne;
    J.flags = LEKF_F_APPLY | (h->infl_mode == LEKF_INFL_FIXED ? LEKF_F_FIXED_INFL : 0) |
              (h->prec == LEKF_PREC_F32 ? LEKF_F_MIXED : 0) |
              (h->prec == LEKF_PREC_BF16X3 ? LEKF_F_BF16X3 : 0);
    J.nij = h->nij;
    J.nv3d = h->nv3d;
    J.nv2d = h->nv2d;
    J.infl_fixed = h->infl_value;
    J.hdxf = h->o_hdxf.as<double>();
    J.y_so = h->ne;
    J.y_sm = 1;
    J.hdxf32 = h->prec != LEKF_PREC_F64 ? h->o_hdxf32.as<float>() : nullptr;
    J.rdiag = h->o_rdiag.as<double>();
    J.dep = h->o_dep.as<double>();
    const bool capped = h->kind == 0 && h->g.nobs_max > 0;  // the solve reads the capped lists
    J.loc_off = (capped ? h->capl_off : h->loc_off).as<long>();
    J.loc_cnt = (capped ? h->capl_cnt : h->loc_cnt).as<int>();
    J.loc_idx = (capped ? h->capl_idx : h->loc_idx).as<int>();
    J.loc_rho = (capped ? h->capl_rho : h->loc_rho).as<double>();
    J.infl_in = h->infl_in.as<double>();
    J.infl_out = h->infl_out.as<double>();
    J.status = h->status.as<int>();
    J.x3b = h->x3b.as<double>();
    J.x3a = h->x3a.as<double>();
    J.x2b = h->nv2d ? h->x2b.as<double>() : nullptr;
    J.x2a = h->nv2d ? h->x2a.as<double>() : nullptr;
    J.upd_mask = h->upd.as<uint32_t>();
    if (h->kind == 2 && h->r.minfl > 0.0) {  // INFL_MUL_MIN (scale/letkf/letkf_tools.f90:199-205)
        J.flags |= LEKF_F_MINFL;
        J.minfl = h->r.minfl;
    }
}

// device job descriptors: one per chunk of points for the regional kind
int upload_jobs(lekf_handle* h)
{
    fill_job(h);
    std::vector<lekf_job> J;
    if (h->kind == 2) {
        for (size_t c = 0; c + 1 < h->chunk_pts.size(); ++c) {
            lekf_job j = h->job;
            j.pt0 = h->chunk_pts[c];
            j.npt = (int)(h->chunk_pts[c + 1] - h->chunk_pts[c]);
            J.push_back(j);
        }
    } else {
        J.push_back(h->job);
    }
    HIPCHK(h->djob.ensure(sizeof(lekf_job) * J.size()));
    HIPCHK(hipMemcpyAsync(h->djob.p, J.data(), sizeof(lekf_job) * J.size(), hipMemcpyHostToDevice,
                          h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

RegionalArgs make_reg_args(const lekf_handle* h)
{
    RegionalArgs R;
    memset(&R, 0, sizeof(R));
    R.nx = h->r.nx;
    R.ny = h->r.ny;
    R.nlev = h->nlev;
    R.nbkt = h->nbkt;
    R.nobs_max = h->r.nobs_max;
    R.dx = h->r.dx;
    R.dy = h->r.dy;
    R.sigma_h = h->r.sigma_h;
    R.sigma_v = h->r.sigma_v;
    R.dlon_zero = h->dlon_zero;
    R.dlat_zero = h->dlat_zero;
    R.zlev = h->d_zlev.as<double>();
    R.cellcum = h->cellcum.as<int>();
    R.ori = h->o_ri.as<double>();
    R.orj = h->o_rj.as<double>();
    R.oz = h->o_z.as<double>();
    R.oerr = h->o_err.as<double>();
    R.obkt = h->o_bkt.as<int>();
    R.bcc = h->b_cc.as<int>();
    R.brec = h->b_rec.as<double>();
    R.bo = h->b_o.as<int>();
    R.err2_min = h->b_err2min.as<unsigned long long>();
    R.wc_so = h->wc_so.as<int>();
    R.wc_so_waves = h->wc_so.p ? lekf_reg_wave_slots_max() : 0;
    R.list_off = h->cap_off.as<long>();
    R.loc_off = h->loc_off.as<long>();
    R.loc_cnt = h->loc_cnt.as<int>();
    R.loc_idx = h->loc_idx.as<int>();
    R.loc_rho = h->loc_rho.as<double>();
    R.s_key = h->sel_key.as<unsigned long long>();
    R.s_ob = h->sel_ob.as<int>();
    R.s_rho = h->sel_rho.as<double>();
    R.slot_cap = h->slot_cap;
    R.g_key = h->sel_gkey.as<unsigned long long>();
    R.g_pos = h->sel_gpos.as<int>();
    R.g_slot = h->sel_gslot.as<int>();
    return R;
}

int reg_select_chunk(lekf_handle* h, int c)
{
    RegionalArgs R = make_reg_args(h);
    R.pt_begin = h->chunk_pts[c];
    R.list_base = h->chunk_off[c];
    HIPCHK(lekf_reg_select_launch(R, h->chunk_pts[c + 1] - h->chunk_pts[c],
                                  std::min(h->sel_slots, lekf_reg_select_slots()), h->st));
    h->held_chunk = c;
    return 0;
}

int run_select(lekf_handle* h)
{
    hipStream_t st = h->st;
    if (h->kind == 2) {
        if (h->chunk_pts.size() != 2) return LEKF_ERR_ARG;  // chunked: lekf_das_run only
        return reg_select_chunk(h, 0);
    }
    if (h->kind == 0) {
        HIPCHK(lekf_logpfm_launch(h->x2b.as<double>(), h->nij, h->nv2d, h->g.ips, h->ne,
                                  h->d_sig.as<double>(), h->nlev, h->d_logpfm.as<double>(), st));
        GlobalCullArgs A;
        memset(&A, 0, sizeof(A));
        A.nij = h->nij;
        A.nlev = h->nlev;
        A.nlon = h->g.nlon;
        A.nlat = h->g.nlat;
        A.lon = h->d_lon.as<double>();
        A.lat = h->d_lat.as<double>();
        A.col_lon = h->d_col_lon.as<double>();
        A.col_lat = h->d_col_lat.as<double>();
        A.col_dlon_zero = h->d_col_dlz.as<double>();
        A.dlat_zero = h->dlat_zero;
        A.dist_zero = h->dist_zero;
        A.dist_zerov = h->dist_zerov;
        A.sigma_obs = h->g.sigma_obs;
        A.sigma_obsv = h->g.sigma_obsv;
        A.logpfm = h->d_logpfm.as<double>();
        A.var_local = h->d_var_local.as<double>();
        A.nobsgrd = h->nobsgrd.as<int>();
        A.o_typ = h->o_typ.as<int>();
        A.o_sinlat = h->o_sin.as<double>();
        A.o_coslat = h->o_cos.as<double>();
        A.o_lonr = h->o_lonr.as<double>();
        A.o_lnlev = h->o_lnlev.as<double>();
        A.o_lndat = h->o_lndat.as<double>();
        A.cand_cnt = h->cand_cnt.as<int>();
        A.cand_off = h->cand_off.as<long>();
        A.loc_off = h->loc_off.as<long>();
        A.loc_cnt = h->loc_cnt.as<int>();
        A.loc_idx = h->loc_idx.as<int>();
        A.loc_rho = h->loc_rho.as<double>();
        if (h->g.nobs_max <= 0) {
            HIPCHK(lekf_cull_global_fill_launch(A, h->nij, st));
        } else {
            // column chunks: raw lists of the chunk, then the nearest-first cap
            ListCapArgs K;
            memset(&K, 0, sizeof(K));
            K.nij = h->nij;
            K.nlev = h->nlev;
            K.cap = h->gcap;
            K.in_off = h->loc_off.as<long>();
            K.in_cnt = h->loc_cnt.as<int>();
            K.in_idx = h->loc_idx.as<int>();
            K.in_rho = h->loc_rho.as<double>();
            K.out_off = h->capl_off.as<long>();
            K.out_cnt = h->capl_cnt.as<int>();
            K.out_idx = h->capl_idx.as<int>();
            K.out_rho = h->capl_rho.as<double>();
            K.g_key = h->sel_gkey.as<unsigned long long>();
            K.g_pos = h->sel_gpos.as<int>();
            K.g_slot = h->sel_gslot.as<int>();
            for (size_t c = 0; c + 1 < h->gchunk_cols.size(); ++c) {
                A.col_begin = h->gchunk_cols[c];
                A.list_base = h->gchunk_base[c];
                const int ncols = h->gchunk_cols[c + 1] - h->gchunk_cols[c];
                HIPCHK(lekf_cull_global_fill_launch(A, ncols, st));
                K.col_begin = A.col_begin;
                K.ncols = ncols;
                HIPCHK(lekf_list_cap_launch(K, std::min(h->sel_slots, lekf_reg_select_slots()), st));
            }
        }
    } else {
        L96CullArgs A;
        A.nx = h->nij;
        A.ny = h->nobs;
        A.gidx = h->r_gidx.as<int>();
        A.xlocal = h->l.xlocal;
        A.xmaxloc = h->l.xlocal * 2.0 * std::sqrt(10.0 / 3.0);
        A.loc_off = h->loc_off.as<long>();
        A.loc_cnt = h->loc_cnt.as<int>();
        A.loc_idx = h->loc_idx.as<int>();
        A.loc_rho = h->loc_rho.as<double>();
        HIPCHK(lekf_cull_l96_launch(A, st));
    }
    return 0;
}

}  // namespace

extern "C" {

lekf_handle* lekf_global_create(const lekf_global_cfg* cfg, int* status)
{
    auto fail = [&](int c) -> lekf_handle* {
        if (status) *status = c;
        return nullptr;
    };
    if (!cfg || cfg->nlon < 2 || cfg->nlat < 3 || cfg->nlev < 1 || cfg->ne < 2 ||
        cfg->ne > LEKF_KMAX || cfg->nv3d < 1 || cfg->nv2d < 1 || cfg->nv3d + cfg->nv2d > LEKF_NVMAX ||
        cfg->nlev > LEKF_MAX_LEV || cfg->ips < 0 || cfg->ips >= cfg->nv2d || cfg->nv3d > 32)
        return fail(LEKF_ERR_SIZE);
    if (lekf_device_count() < 1) return fail(LEKF_ERR_NODEV);
    lekf_handle* h = new lekf_handle();
    h->kind = 0;
    h->g = *cfg;
    h->ne = cfg->ne;
    h->nij = cfg->nlon * cfg->nlat;
    h->nlev = cfg->nlev;
    h->nv3d = cfg->nv3d;
    h->nv2d = cfg->nv2d;
    h->prec = cfg->prec;
    h->infl_mode = cfg->infl_mode;
    h->infl_value = cfg->infl_value;
    h->lon.assign(cfg->lon, cfg->lon + cfg->nlon);
    h->lat.assign(cfg->lat, cfg->lat + cfg->nlat);
    h->sig.assign(cfg->sig, cfg->sig + cfg->nlev);
    h->g.lon = h->lon.data();
    h->g.lat = h->lat.data();
    h->g.sig = h->sig.data();
    int rc = 0;
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(LEKF_ERR_HIP);
    }
    // localisation constants (speedy/letkf/letkf_obs.f90:82-88); host libm as in the reference
    const double pi = LEKF_PI_REF, re = LEKF_RE_REF;
    h->dist_zero = cfg->sigma_obs * std::sqrt(10.0 / 3.0) * 2.0;
    h->dist_zerov = cfg->sigma_obsv * std::sqrt(10.0 / 3.0) * 2.0;
    h->dlat_zero = h->dist_zero / pi / re * 180.0;
    std::vector<double> clon(h->nij), clat(h->nij), cdz(h->nij);
    for (int j = 0; j < cfg->nlat; ++j)
        for (int i = 0; i < cfg->nlon; ++i) {
            const int ij = i + cfg->nlon * j;
            clon[ij] = cfg->lon[i];
            clat[ij] = cfg->lat[j];
            cdz[ij] = h->dlat_zero / std::cos(pi * cfg->lat[j] / 180.0);
        }
    hipStream_t st = h->st;
    if ((rc = upload(h->d_lon, cfg->lon, sizeof(double) * cfg->nlon, st)) ||
        (rc = upload(h->d_lat, cfg->lat, sizeof(double) * cfg->nlat, st)) ||
        (rc = upload(h->d_sig, cfg->sig, sizeof(double) * cfg->nlev, st)) ||
        (rc = upload(h->d_col_lon, clon.data(), sizeof(double) * h->nij, st)) ||
        (rc = upload(h->d_col_lat, clat.data(), sizeof(double) * h->nij, st)) ||
        (rc = upload(h->d_col_dlz, cdz.data(), sizeof(double) * h->nij, st)) ||
        (rc = upload(h->d_var_local, cfg->var_local, sizeof(double) * 7, st)) ||
        (rc = (h->d_logpfm.ensure(sizeof(double) * (size_t)h->nij * h->nlev) == hipSuccess
                   ? 0
                   : LEKF_ERR_HIP)) ||
        (rc = (h->cand_cnt.ensure(sizeof(int) * h->nij) == hipSuccess ? 0 : LEKF_ERR_HIP)) ||
        (rc = (h->cand_off.ensure(sizeof(long) * (h->nij + 1)) == hipSuccess ? 0 : LEKF_ERR_HIP)) ||
        (rc = alloc_state(h))) {
        free_all(h);
        (void)hipStreamDestroy(h->st);
        delete h;
        return fail(rc);
    }
    if (status) *status = 0;
    return h;
}

lekf_handle* lekf_l96_create(const lekf_l96_cfg* cfg, int* status)
{
    auto fail = [&](int c) -> lekf_handle* {
        if (status) *status = c;
        return nullptr;
    };
    if (!cfg || cfg->nx < 1 || cfg->ne < 2 || cfg->ne > LEKF_KMAX) return fail(LEKF_ERR_SIZE);
    if (lekf_device_count() < 1) return fail(LEKF_ERR_NODEV);
    lekf_handle* h = new lekf_handle();
    h->kind = 1;
    h->l = *cfg;
    h->ne = cfg->ne;
    h->nij = cfg->nx;
    h->nlev = 1;
    h->nv3d = 1;
    h->nv2d = 0;
    h->prec = cfg->prec;
    h->infl_mode = cfg->infl_mode;
    h->infl_value = cfg->infl_value;
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(LEKF_ERR_HIP);
    }
    int rc = alloc_state(h);
    if (rc) {
        free_all(h);
        (void)hipStreamDestroy(h->st);
        delete h;
        return fail(rc);
    }
    if (status) *status = 0;
    return h;
}

void graph_drop(lekf_handle* h)
{
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
    h->graph_ok = false;
}

void lekf_destroy(lekf_handle* h)
{
    if (!h) return;
    (void)hipStreamSynchronize(h->st);
    graph_drop(h);
    free_all(h);
    if (h->own_stream) (void)hipStreamDestroy(h->st);
    delete h;
}

void* lekf_stream(lekf_handle* h) { return h ? (void*)h->st : nullptr; }

long lekf_device_bytes(lekf_handle* h) { return h ? handle_bytes(h) : 0; }

int lekf_set_ensemble(lekf_handle* h, const double* gues3d, const double* gues2d)
{
    if (!h || !gues3d || (h->nv2d > 0 && !gues2d)) return LEKF_ERR_ARG;
    if (h) graph_drop(h);
    hipStream_t st = h->st;
    const size_t n3 = (size_t)h->nij * h->nlev * h->ne * h->nv3d;
    const size_t n2 = (size_t)h->nij * h->ne * h->nv2d;
    HIPCHK(h->stage.ensure(sizeof(double) * std::max(n3, n2)));
    HIPCHK(hipMemcpyAsync(h->stage.p, gues3d, sizeof(double) * n3, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(ens3_to_dev, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st,
                       h->stage.as<double>(), h->x3b.as<double>(), h->nij, h->nlev, h->ne, h->nv3d);
    HIPCHK(hipGetLastError());
    if (n2) {
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipMemcpyAsync(h->stage.p, gues2d, sizeof(double) * n2, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(ens3_to_dev, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st,
                           h->stage.as<double>(), h->x2b.as<double>(), h->nij, 1, h->ne, h->nv2d);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(st));
    h->have_ens = true;
    return LEKF_OK;
}

int lekf_set_infl(lekf_handle* h, const double* infl3d, const double* infl2d)
{
    if (!h || !infl3d) return LEKF_ERR_ARG;
    if (h) graph_drop(h);
    const size_t npt = (size_t)h->nij * h->nlev;
    // reference infl3d(nij,nlev) column-major == device point order lev*nij + ij
    HIPCHK(hipMemcpyAsync(h->infl_in.p, infl3d, sizeof(double) * npt, hipMemcpyHostToDevice, h->st));
    if (infl2d && h->nv2d)
        HIPCHK(hipMemcpyAsync(h->infl2.p, infl2d, sizeof(double) * h->nij * h->nv2d,
                              hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return LEKF_OK;
}

int upload_obs_records(lekf_handle* h, int nobs, const double* elm, const double* lon,
                       const double* lat, const double* lev, const double* dat, const double* err)
{
    hipStream_t st = h->st;
    int rc;
    if ((rc = upload(h->r_elm, elm, sizeof(double) * nobs, st)) ||
        (rc = upload(h->r_lon, lon, sizeof(double) * nobs, st)) ||
        (rc = upload(h->r_lat, lat, sizeof(double) * nobs, st)) ||
        (rc = upload(h->r_lev, lev, sizeof(double) * nobs, st)) ||
        (rc = upload(h->r_dat, dat, sizeof(double) * nobs, st)))
        return rc;
    if (err && (rc = upload(h->r_err, err, sizeof(double) * nobs, st))) return rc;
    return 0;
}

// H(x) of the resident background into r_hxf (row = obs, k members) and the
// member-combined QC flag into r_qc
int run_obsope_global(lekf_handle* h, int nobs, double* ri, double* rj, double* rk)
{
    hipStream_t st = h->st;
    const size_t n = (size_t)(nobs > 0 ? nobs : 1);
    HIPCHK(h->r_hxf.ensure(sizeof(double) * n * h->ne));
    HIPCHK(h->r_qc.ensure(sizeof(int) * n));
    ObsOpArgs O;
    memset(&O, 0, sizeof(O));
    O.nobs = nobs;
    O.ne = h->ne;
    O.nlon = h->g.nlon;
    O.nlat = h->g.nlat;
    O.nlev = h->nlev;
    O.nv3d = h->nv3d;
    O.nv2d = h->nv2d;
    O.iu = 0;
    O.iv = 1;
    O.it = 2;
    O.iq = h->g.iq >= 0 ? h->g.iq : 3;
    O.ips = h->g.ips;
    O.lat = h->d_lat.as<double>();
    O.sig = h->d_sig.as<double>();
    O.x3 = h->x3b.as<double>();
    O.x2 = h->x2b.as<double>();
    O.elm = h->r_elm.as<double>();
    O.olon = h->r_lon.as<double>();
    O.olat = h->r_lat.as<double>();
    O.olev = h->r_lev.as<double>();
    O.odat = h->r_dat.as<double>();
    O.hxf = h->r_hxf.as<double>();
    O.qc = h->r_qc.as<int>();
    O.ri = ri;
    O.rj = rj;
    O.rk = rk;
    HIPCHK(lekf_obsope_global_launch(O, st));
    return 0;
}

int set_obs_global_dev(lekf_handle* h, int nobs, bool have_qc, long h_so, long h_sm);

int lekf_set_obs_global(lekf_handle* h, int nobs, const double* elm, const double* lon,
                        const double* lat, const double* lev, const double* dat,
                        const double* err, const int* qc, const double* hxf)
{
    if (!h || h->kind != 0 || nobs < 0) return LEKF_ERR_ARG;
    if (h) graph_drop(h);
    if (nobs > 0 && (!elm || !lon || !lat || !lev || !dat || !err || !hxf)) return LEKF_ERR_ARG;
    hipStream_t st = h->st;
    int rc = upload_obs_records(h, nobs, elm, lon, lat, lev, dat, err);
    if (!rc) rc = upload(h->r_hxf, hxf, sizeof(double) * nobs * (size_t)h->ne, st);
    if (rc) return rc;
    if (qc && (rc = upload(h->r_qc, qc, sizeof(int) * nobs, st))) return rc;
    return set_obs_global_dev(h, nobs, qc != nullptr, 1, nobs);
}

int lekf_set_obs_global_hx(lekf_handle* h, int nobs, const double* elm, const double* lon,
                           const double* lat, const double* lev, const double* dat,
                           const double* err)
{
    if (!h || h->kind != 0 || nobs < 0 || !h->have_ens) return LEKF_ERR_ARG;
    if (h) graph_drop(h);
    if (nobs > 0 && (!elm || !lon || !lat || !lev || !dat || !err)) return LEKF_ERR_ARG;
    int rc = upload_obs_records(h, nobs, elm, lon, lat, lev, dat, err);
    if (!rc) rc = run_obsope_global(h, nobs, nullptr, nullptr, nullptr);
    if (rc) return rc;
    return set_obs_global_dev(h, nobs, true, h->ne, 1);
}

int lekf_obsope_global(lekf_handle* h, int nobs, const double* elm, const double* lon,
                       const double* lat, const double* lev, const double* dat, double* hxf,
                       int* qc, double* ri, double* rj, double* rk)
{
    if (!h || h->kind != 0 || nobs < 0 || !h->have_ens || !hxf || !qc) return LEKF_ERR_ARG;
    if (nobs == 0) return LEKF_OK;
    if (!elm || !lon || !lat || !lev || !dat) return LEKF_ERR_ARG;
    hipStream_t st = h->st;
    int rc = upload_obs_records(h, nobs, elm, lon, lat, lev, dat, nullptr);
    if (rc) return rc;
    DevBuf dri, drj, drk;
    HIPCHK(dri.ensure(sizeof(double) * nobs));
    HIPCHK(drj.ensure(sizeof(double) * nobs));
    HIPCHK(drk.ensure(sizeof(double) * nobs * (size_t)h->ne));
    rc = run_obsope_global(h, nobs, dri.as<double>(), drj.as<double>(), drk.as<double>());
    if (rc) return rc;
    const size_t ne = (size_t)h->ne;
    std::vector<double> t((size_t)nobs * ne);
    HIPCHK(hipMemcpyAsync(t.data(), h->r_hxf.p, sizeof(double) * t.size(), hipMemcpyDeviceToHost,
                          st));
    HIPCHK(hipMemcpyAsync(qc, h->r_qc.p, sizeof(int) * nobs, hipMemcpyDeviceToHost, st));
    if (ri) HIPCHK(hipMemcpyAsync(ri, dri.p, sizeof(double) * nobs, hipMemcpyDeviceToHost, st));
    if (rj) HIPCHK(hipMemcpyAsync(rj, drj.p, sizeof(double) * nobs, hipMemcpyDeviceToHost, st));
    std::vector<double> tk(rk ? t.size() : 0);
    if (rk) HIPCHK(hipMemcpyAsync(tk.data(), drk.p, sizeof(double) * tk.size(),
                                  hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    // device rows (obs, members) -> reference column-major (nobs, ne)
    for (size_t n = 0; n < (size_t)nobs; ++n)
        for (size_t m = 0; m < ne; ++m) {
            hxf[n + m * nobs] = t[n * ne + m];
            if (rk) rk[n + m * nobs] = tk[n * ne + m];
        }
    dri.release();
    drj.release();
    drk.release();
    return LEKF_OK;
}

int lekf_set_heights(lekf_handle* h, const double* hgt)
{
    if (!h || h->kind != 2 || !hgt) return LEKF_ERR_ARG;
    const int rc = upload(h->d_hgt, hgt, sizeof(double) * (size_t)h->nij * h->nlev, h->st);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->st));
    return LEKF_OK;
}

int lekf_obs_departure_scale(lekf_handle* h, const lekf_scale_qc_cfg* cfg, int nobs,
                             const double* elm, const double* dat, const double* err,
                             const int* qc_member, double* hx, double* dat_out, double* err_out,
                             double* dep, int* qc)
{
    if (!h || !cfg || nobs < 0 || h->ne < 1) return LEKF_ERR_ARG;
    if (nobs == 0) return LEKF_OK;
    if (!elm || !dat || !err || !hx || !dat_out || !err_out || !dep || !qc) return LEKF_ERR_ARG;
    for (int n = 0; n < nobs; ++n)
        if ((int)elm[n] == LEKF_ID_H08IR) return LEKF_ERR_ARG;
    hipStream_t st = h->st;
    const size_t n = (size_t)nobs, ne = (size_t)h->ne;
    DevBuf rec, dhx, dqm, dq;
    struct Release {  // the call's buffers go on every return path
        DevBuf* b[4];
        ~Release()
        {
            for (DevBuf* x : b) x->release();
        }
    } rel{{&rec, &dhx, &dqm, &dq}};
    HIPCHK(rec.ensure(sizeof(double) * n * 6));
    const double* in[3] = {elm, dat, err};
    for (int f = 0; f < 3; ++f)
        HIPCHK(hipMemcpyAsync(rec.as<double>() + n * f, in[f], sizeof(double) * n,
                              hipMemcpyHostToDevice, st));
    HIPCHK(dhx.ensure(sizeof(double) * n * ne));
    HIPCHK(hipMemcpyAsync(dhx.p, hx, sizeof(double) * n * ne, hipMemcpyHostToDevice, st));
    if (qc_member) {
        HIPCHK(dqm.ensure(sizeof(int) * n * ne));
        HIPCHK(hipMemcpyAsync(dqm.p, qc_member, sizeof(int) * n * ne, hipMemcpyHostToDevice, st));
    }
    HIPCHK(dq.ensure(sizeof(int) * n));
    auto factor = [&](double f) { return f < 0.0 ? cfg->gross_error : f; };
    ScaleQcArgs A{};
    A.nobs = nobs;
    A.ne = h->ne;
    A.use_ref = cfg->use_radar_ref != 0;
    A.use_vr = cfg->use_radar_vr != 0;
    A.use_obserr_ref = cfg->use_obserr_radar_ref != 0;
    A.use_obserr_vr = cfg->use_obserr_radar_vr != 0;
    A.min_member = cfg->min_radar_ref_member;
    A.min_member_obsref = cfg->min_radar_ref_member_obsref;
    // common_nml.f90:596-598 and common_obs_scale.f90:249
    A.thres = std::max(cfg->radar_ref_thres_dbz, cfg->min_radar_ref_dbz);
    A.min_ref_dbz = cfg->min_radar_ref_dbz;
    A.low_ref_shift = cfg->low_ref_shift;
    A.min_ref = std::pow(10.0, cfg->min_radar_ref_dbz / 10.0);
    A.obserr_ref = cfg->obserr_radar_ref;
    A.obserr_vr = cfg->obserr_radar_vr;
    A.gross = cfg->gross_error;
    A.gross_rain = factor(cfg->gross_error_rain);
    A.gross_ref = factor(cfg->gross_error_radar_ref);
    A.gross_vr = factor(cfg->gross_error_radar_vr);
    A.gross_prh = factor(cfg->gross_error_radar_prh);
    A.gross_tcx = factor(cfg->gross_error_tcx);
    A.gross_tcy = factor(cfg->gross_error_tcy);
    A.gross_tcp = factor(cfg->gross_error_tcp);
    double* r = rec.as<double>();
    A.elm = r;
    A.dat_in = r + n;
    A.err_in = r + 2 * n;
    A.dat = r + 3 * n;
    A.err = r + 4 * n;
    A.dep = r + 5 * n;
    A.so = 1;
    A.sm = nobs;
    A.qc_mem = qc_member ? dqm.as<int>() : nullptr;
    A.hx = dhx.as<double>();
    A.qc = dq.as<int>();
    HIPCHK(lekf_scale_qc_launch(A, st));
    HIPCHK(hipMemcpyAsync(dat_out, r + 3 * n, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(err_out, r + 4 * n, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(dep, r + 5 * n, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hx, dhx.p, sizeof(double) * n * ne, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(qc, dq.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return LEKF_OK;
}

int lekf_obsope_radar(lekf_handle* h, const lekf_radar_cfg* cfg, int nobs, const double* elm,
                      const double* lon, const double* lat, const double* z, const double* ri,
                      const double* rj, const double* radar_lon, const double* radar_lat,
                      const double* radar_z, double* hxf, int* qc, double* rk)
{
    if (!h || !cfg || h->kind != 2 || nobs < 0 || !h->have_ens || !h->d_hgt.p || !hxf || !qc ||
        h->nlev < 2)
        return LEKF_ERR_ARG;
    if (cfg->method < 1 || cfg->method > 3) return LEKF_ERR_ARG;
    for (int q = 0; q < 11; ++q)
        if (cfg->var[q] < 0 || cfg->var[q] >= h->nv3d) return LEKF_ERR_ARG;
    if (nobs == 0) return LEKF_OK;
    const double* in[9] = {elm, lon, lat, z, ri, rj, radar_lon, radar_lat, radar_z};
    for (auto* p : in)
        if (!p) return LEKF_ERR_ARG;
    hipStream_t st = h->st;
    const size_t n = (size_t)nobs, ne = (size_t)h->ne;
    DevBuf rec, dhx, dqc, drk;
    struct Release {  // the call's buffers go on every return path
        DevBuf* b[4];
        ~Release()
        {
            for (DevBuf* x : b) x->release();
        }
    } rel{{&rec, &dhx, &dqc, &drk}};
    HIPCHK(rec.ensure(sizeof(double) * n * 9));
    for (int f = 0; f < 9; ++f)
        HIPCHK(hipMemcpyAsync(rec.as<double>() + n * f, in[f], sizeof(double) * n,
                              hipMemcpyHostToDevice, st));
    HIPCHK(dhx.ensure(sizeof(double) * n * ne));
    HIPCHK(dqc.ensure(sizeof(int) * n * ne));
    HIPCHK(drk.ensure(sizeof(double) * n));
    RadarArgs A{};
    A.nobs = nobs;
    A.nx = h->r.nx;
    A.ny = h->r.ny;
    A.nlev = h->nlev;
    A.ne = h->ne;
    A.nv3d = h->nv3d;
    for (int q = 0; q < 11; ++q) A.var[q] = cfg->var[q];
    A.method = cfg->method;
    A.use_wt = cfg->use_terminal_velocity ? 1 : 0;
    A.min_ref_dbz = cfg->min_ref_dbz;
    A.low_ref_shift = cfg->low_ref_shift;
    A.zmax = cfg->zmax;
    // com_gamma of 4 + b, 4 + d and 4.5 in calc_ref_vr
    A.gam[0] = std::tgamma(4.8);
    A.gam[1] = std::tgamma(4.25);
    A.gam[2] = std::tgamma(4.5);
    const double* r = rec.as<double>();
    A.elm = r;
    A.lon = r + n;
    A.lat = r + 2 * n;
    A.z = r + 3 * n;
    A.ri = r + 4 * n;
    A.rj = r + 5 * n;
    A.rlon = r + 6 * n;
    A.rlat = r + 7 * n;
    A.rz = r + 8 * n;
    A.x3 = h->x3b.as<double>();
    A.hgt = h->d_hgt.as<double>();
    A.hxf = dhx.as<double>();
    A.qc = dqc.as<int>();
    A.rk = drk.as<double>();
    HIPCHK(lekf_radar_launch(A, st));
    std::vector<double> t(n * ne);
    std::vector<int> tq(n * ne);
    HIPCHK(hipMemcpyAsync(t.data(), dhx.p, sizeof(double) * t.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(tq.data(), dqc.p, sizeof(int) * tq.size(), hipMemcpyDeviceToHost, st));
    if (rk) HIPCHK(hipMemcpyAsync(rk, drk.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    // device rows (obs, members) -> column-major (nobs, ne)
    for (size_t i = 0; i < n; ++i)
        for (size_t m = 0; m < ne; ++m) {
            hxf[i + m * n] = t[i * ne + m];
            qc[i + m * n] = tq[i * ne + m];
        }
    return LEKF_OK;
}

// departures, QC and binning of obs records and H(x) already on the device
// (r_elm .. r_err, r_hxf with member stride h_sm and obs stride h_so, r_qc)
int set_obs_global_dev(lekf_handle* h, int nobs, bool have_qc, long h_so, long h_sm)
{
    hipStream_t st = h->st;
    const int ne = h->ne;
    const size_t n = (size_t)(nobs > 0 ? nobs : 1);
    long n2 = 2048;
    while (n2 < nobs) n2 <<= 1;
    const int ncell = h->g.nlon * h->g.nlat;
    HIPCHK(h->keys.ensure(sizeof(unsigned long long) * n2));
    HIPCHK(h->cell_cnt.ensure(sizeof(int) * ncell));
    HIPCHK(h->nobsgrd.ensure(sizeof(int) * ncell));
    HIPCHK(h->tmp_mean.ensure(sizeof(double) * n));
    HIPCHK(h->tmp_dep.ensure(sizeof(double) * n));
    HIPCHK(h->perm.ensure(sizeof(int) * n));
    HIPCHK(h->o_hdxf.ensure(sizeof(double) * n * ne));
    if (h->prec != LEKF_PREC_F64) HIPCHK(h->o_hdxf32.ensure(sizeof(float) * n * ne));
    HIPCHK(h->o_dep.ensure(sizeof(double) * n));
    HIPCHK(h->o_rdiag.ensure(sizeof(double) * n));
    HIPCHK(h->o_err.ensure(sizeof(double) * n));
    HIPCHK(h->o_typ.ensure(sizeof(int) * n));
    HIPCHK(h->o_sin.ensure(sizeof(double) * n));
    HIPCHK(h->o_cos.ensure(sizeof(double) * n));
    HIPCHK(h->o_lonr.ensure(sizeof(double) * n));
    HIPCHK(h->o_lnlev.ensure(sizeof(double) * n));
    HIPCHK(h->o_lndat.ensure(sizeof(double) * n));
    HIPCHK(hipMemsetAsync(h->cell_cnt.p, 0, sizeof(int) * ncell, st));
    HIPCHK(hipMemsetAsync(h->keys.p, 0xff, sizeof(unsigned long long) * n2, st));
    ObsPrepArgs A;
    memset(&A, 0, sizeof(A));
    A.nobs = nobs;
    A.ne = ne;
    A.nlon = h->g.nlon;
    A.nlat = h->g.nlat;
    A.lon = h->d_lon.as<double>();
    A.lat = h->d_lat.as<double>();
    A.elm = h->r_elm.as<double>();
    A.olon = h->r_lon.as<double>();
    A.olat = h->r_lat.as<double>();
    A.olev = h->r_lev.as<double>();
    A.dat = h->r_dat.as<double>();
    A.err = h->r_err.as<double>();
    A.qc = have_qc ? h->r_qc.as<int>() : nullptr;
    A.hxf = h->r_hxf.as<double>();
    A.h_so = h_so;
    A.h_sm = h_sm;
    A.gross_error = h->g.gross_error;
    A.tmp_mean = h->tmp_mean.as<double>();
    A.tmp_dep = h->tmp_dep.as<double>();
    A.keys = h->keys.as<unsigned long long>();
    A.cell_cnt = h->cell_cnt.as<int>();
    A.perm = h->perm.as<int>();
    A.o_hdxf = h->o_hdxf.as<double>();
    A.o_hdxf32 = h->prec != LEKF_PREC_F64 ? h->o_hdxf32.as<float>() : nullptr;
    A.o_dep = h->o_dep.as<double>();
    A.o_rdiag = h->o_rdiag.as<double>();
    A.o_err = h->o_err.as<double>();
    A.o_typ = h->o_typ.as<int>();
    A.o_sinlat = h->o_sin.as<double>();
    A.o_coslat = h->o_cos.as<double>();
    A.o_lonr = h->o_lonr.as<double>();
    A.o_lnlev = h->o_lnlev.as<double>();
    A.o_lndat = h->o_lndat.as<double>();
    A.nobsgrd = h->nobsgrd.as<int>();
    if (nobs > 0) HIPCHK(lekf_obs_prep_launch(A, st));
    HIPCHK(lekf_sort_keys_launch(A.keys, n2, st));
    // binned count = total of the cell histogram
    std::vector<int> cc(ncell);
    HIPCHK(hipMemcpyAsync(cc.data(), h->cell_cnt.p, sizeof(int) * ncell, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    long nb = 0;
    for (int c : cc) nb += c;
    h->nbinned = (int)nb;
    h->nobs = nobs;
    HIPCHK(lekf_obs_scatter_launch(A, h->nbinned, ncell, st));
    // candidate capacity per column (depends on obs geometry only)
    GlobalCullArgs C;
    memset(&C, 0, sizeof(C));
    C.nij = h->nij;
    C.nlev = h->nlev;
    C.nlon = h->g.nlon;
    C.nlat = h->g.nlat;
    C.lon = h->d_lon.as<double>();
    C.lat = h->d_lat.as<double>();
    C.col_lon = h->d_col_lon.as<double>();
    C.col_lat = h->d_col_lat.as<double>();
    C.col_dlon_zero = h->d_col_dlz.as<double>();
    C.dlat_zero = h->dlat_zero;
    C.nobsgrd = h->nobsgrd.as<int>();
    C.cand_cnt = h->cand_cnt.as<int>();
    C.cand_off = h->cand_off.as<long>();
    HIPCHK(lekf_cull_global_count_launch(C, st));
    long tot = 0;
    HIPCHK(hipMemcpyAsync(&tot, h->cand_off.as<long>() + h->nij, sizeof(long), hipMemcpyDeviceToHost,
                          st));
    HIPCHK(hipStreamSynchronize(st));
    if (h->g.nobs_max <= 0) {
        h->loc_cap = tot * h->nlev;
    } else {
        // raw lists in column chunks that fit the list budget; capped lists for all points
        std::vector<long> co(h->nij + 1);
        HIPCHK(hipMemcpy(co.data(), h->cand_off.p, sizeof(long) * (h->nij + 1),
                         hipMemcpyDeviceToHost));
        h->gchunk_cols.assign(1, 0);
        h->gchunk_base.assign(1, 0);
        long maxraw = 1, maxcol = 1;
        for (int c = 0; c < h->nij;) {
            int d = c + 1;
            while (d < h->nij && (co[d + 1] - co[c]) * h->nlev <= h->list_budget) ++d;
            h->gchunk_cols.push_back(d);
            h->gchunk_base.push_back(co[d] * h->nlev);
            maxraw = std::max(maxraw, (co[d] - co[c]) * h->nlev);
            c = d;
        }
        for (int c = 0; c < h->nij; ++c) maxcol = std::max(maxcol, co[c + 1] - co[c]);
        h->loc_cap = maxraw;
        // no list is longer than its column's candidate count: a larger cap
        // never binds, and the capped lists need no more room than that
        h->gcap = (int)std::min<long>(h->g.nobs_max, maxcol);
        const size_t npt = (size_t)h->nij * h->nlev, cap = (size_t)h->gcap;
        HIPCHK(h->capl_off.ensure(sizeof(long) * npt));
        HIPCHK(h->capl_cnt.ensure(sizeof(int) * npt));
        HIPCHK(h->capl_idx.ensure(sizeof(int) * npt * cap));
        HIPCHK(h->capl_rho.ensure(sizeof(double) * npt * cap));
        h->sel_slots = lekf_reg_select_slots_max();
        HIPCHK(ensure_cap_sets(h, (int)cap));
    }
    HIPCHK(h->loc_idx.ensure(sizeof(int) * (size_t)(h->loc_cap ? h->loc_cap : 1)));
    HIPCHK(h->loc_rho.ensure(sizeof(double) * (size_t)(h->loc_cap ? h->loc_cap : 1)));
    h->have_obs = true;
    fill_job(h);
    HIPCHK(h->djob.ensure(sizeof(lekf_job)));
    HIPCHK(hipMemcpyAsync(h->djob.p, &h->job, sizeof(lekf_job), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return LEKF_OK;
}

int lekf_set_obs_l96(lekf_handle* h, int ny, const int* gidx, const double* y, const double* err,
                     const double* hxf)
{
    if (!h || h->kind != 1 || ny < 0) return LEKF_ERR_ARG;
    if (!hxf && !h->have_ens) return LEKF_ERR_ARG;
    for (int n = 0; n < ny; ++n)
        if (gidx[n] < 0 || gidx[n] >= h->nij) return LEKF_ERR_ARG;
    if (h) graph_drop(h);
    hipStream_t st = h->st;
    const int ne = h->ne;
    const size_t n = (size_t)(ny > 0 ? ny : 1);
    int rc;
    if ((rc = upload(h->r_gidx, gidx, sizeof(int) * ny, st)) ||
        (rc = upload(h->r_y, y, sizeof(double) * ny, st)) ||
        (rc = upload(h->r_err, err, sizeof(double) * ny, st)))
        return rc;
    if (hxf) {
        if ((rc = upload(h->r_hxf, hxf, sizeof(double) * ny * (size_t)ne, st))) return rc;
    } else {
        // H(x) of the resident background (lorenz96/obs/h_ope.f90:14-39)
        HIPCHK(h->r_hxf.ensure(sizeof(double) * n * ne));
        HIPCHK(lekf_l96_hx_launch(ny, ne, h->r_gidx.as<int>(), h->x3b.as<double>(),
                                  h->r_hxf.as<double>(), st));
    }
    HIPCHK(h->o_hdxf.ensure(sizeof(double) * n * ne));
    if (h->prec != LEKF_PREC_F64) HIPCHK(h->o_hdxf32.ensure(sizeof(float) * n * ne));
    HIPCHK(h->o_dep.ensure(sizeof(double) * n));
    HIPCHK(h->o_rdiag.ensure(sizeof(double) * n));
    if (ny > 0) {
        hipLaunchKernelGGL(l96_obs_kernel, dim3((ny + 255) / 256), dim3(256), 0, st, ny, ne,
                           h->r_hxf.as<double>(), h->r_y.as<double>(), h->r_err.as<double>(),
                           h->o_hdxf.as<double>(),
                           h->prec != LEKF_PREC_F64 ? h->o_hdxf32.as<float>() : nullptr,
                           h->o_dep.as<double>(), h->o_rdiag.as<double>());
        HIPCHK(hipGetLastError());
    }
    h->nobs = ny;
    h->nbinned = ny;
    h->loc_cap = (long)h->nij * (ny > 0 ? ny : 1);
    HIPCHK(h->loc_idx.ensure(sizeof(int) * h->loc_cap));
    HIPCHK(h->loc_rho.ensure(sizeof(double) * h->loc_cap));
    h->have_obs = true;
    fill_job(h);
    HIPCHK(h->djob.ensure(sizeof(lekf_job)));
    HIPCHK(hipMemcpyAsync(h->djob.p, &h->job, sizeof(lekf_job), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return LEKF_OK;
}

int lekf_das_select(lekf_handle* h)
{
    if (!h || !h->have_ens || !h->have_obs) return LEKF_ERR_ARG;
    return run_select(h);
}

static bool relax_active(const lekf_handle* h)
{
    return h->relax_mode != LEKF_RELAX_NONE || h->relax_qmax > 0.0 || h->relax_width > 0.0;
}

// the relaxation pass over every point's analysis rows (after all solves)
static int run_relax(lekf_handle* h)
{
    if (!relax_active(h)) return LEKF_OK;
    RelaxArgs A{};
    A.npt = (long)h->nij * h->nlev;
    A.nij = h->nij;
    A.ne = h->ne;
    A.nv3d = h->nv3d;
    A.nv2d = h->nv2d;
    A.mode = h->relax_mode;
    A.alpha = h->relax_alpha;
    A.qmax = h->relax_qmax;
    A.iq = h->relax_iq;
    A.width = h->relax_width;
    if (h->kind == 2) {
        A.nx = h->r.nx;
        A.ny = h->r.ny;
        A.dx = h->r.dx;
        A.dy = h->r.dy;
    }
    A.x3b = h->x3b.as<double>();
    A.x3a = h->x3a.as<double>();
    A.x2b = h->nv2d ? h->x2b.as<double>() : nullptr;
    A.x2a = h->nv2d ? h->x2a.as<double>() : nullptr;
    HIPCHK(lekf_relax_launch(A, h->st));
    return LEKF_OK;
}

int lekf_das_solve(lekf_handle* h)
{
    if (!h || !h->have_ens || !h->have_obs) return LEKF_ERR_ARG;
    long npt = h->job.npt;
    if (h->kind == 2) {
        if (h->chunk_pts.size() != 2) return LEKF_ERR_ARG;
        npt = h->chunk_pts[1];
    }
    const int rc = run_solve(h->djob.as<lekf_job>(), 1, npt, h->ne, job_rows(h->job),
                             mode_of_prec(h->prec), h->gws, h->st);
    if (rc) return rc;
    return run_relax(h);
}

int das_run_eager(lekf_handle* h)
{
    if (h && h->kind == 2 && h->have_ens && h->have_obs) {
        // chunks of points whose local lists fit the list buffer: select, then solve
        for (size_t c = 0; c + 1 < h->chunk_pts.size(); ++c) {
            int rc = reg_select_chunk(h, (int)c);
            if (rc) return rc;
            rc = run_solve(h->djob.as<lekf_job>() + c, 1, h->chunk_pts[c + 1] - h->chunk_pts[c],
                           h->ne, job_rows(h->job), mode_of_prec(h->prec), h->gws, h->st);
            if (rc) return rc;
        }
        return run_relax(h);
    }
    int rc = lekf_das_select(h);
    if (rc) return rc;
    return lekf_das_solve(h);
}

int lekf_das_run(lekf_handle* h)
{
    if (!h) return LEKF_ERR_ARG;
    if (!h->use_graph) return das_run_eager(h);
    if (h->graph_ok && h->graph_epoch != g_tune_epoch.load()) {
        HIPCHK(hipStreamSynchronize(h->st));
        graph_drop(h);
    }
    if (h->graph_ok) {
        HIPCHK(hipGraphLaunch(h->gexec, h->st));
        return LEKF_OK;
    }
    // first cycle after a change: run eagerly (this also sizes every work
    // buffer), then capture the same launch sequence for the next cycles
    const long epoch = g_tune_epoch.load();
    int rc = das_run_eager(h);
    if (rc) return rc;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
    rc = das_run_eager(h);
    const hipError_t e = hipStreamEndCapture(h->st, &g);
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (!rc) set_err(e, "hipStreamEndCapture");
        return rc ? rc : LEKF_ERR_HIP;
    }
    const hipError_t ei = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(ei);
    h->graph_ok = true;
    h->graph_epoch = epoch;
    return LEKF_OK;
}

int lekf_set_relax(lekf_handle* h, int mode, double alpha, double q_sprd_max, int iq,
                   double boundary_width)
{
    if (!h) return LEKF_ERR_ARG;
    if (mode < LEKF_RELAX_NONE || mode > LEKF_RELAX_RTPS || !(alpha >= 0.0) ||
        !(q_sprd_max >= 0.0) || !(boundary_width >= 0.0))
        return LEKF_ERR_ARG;
    if (q_sprd_max > 0.0 && (iq < 0 || iq >= h->nv3d)) return LEKF_ERR_ARG;
    if (boundary_width > 0.0 && h->kind != 2) return LEKF_ERR_ARG;
    HIPCHK(hipStreamSynchronize(h->st));
    graph_drop(h);
    h->relax_mode = mode;
    h->relax_alpha = alpha;
    h->relax_qmax = q_sprd_max;
    h->relax_iq = q_sprd_max > 0.0 ? iq : -1;
    h->relax_width = boundary_width;
    return LEKF_OK;
}

int lekf_set_graph(lekf_handle* h, int enable)
{
    if (!h) return LEKF_ERR_ARG;
    HIPCHK(hipStreamSynchronize(h->st));
    graph_drop(h);
    h->use_graph = enable != 0;
    return LEKF_OK;
}

lekf_handle* lekf_regional_create(const lekf_regional_cfg* cfg, int* status)
{
    auto fail = [&](int c) -> lekf_handle* {
        if (status) *status = c;
        return nullptr;
    };
    if (!cfg || cfg->nx < 1 || cfg->ny < 1 || cfg->nlev < 1 || cfg->ne < 2 || cfg->ne > LEKF_KMAX ||
        cfg->nv3d < 1 || cfg->nv3d > LEKF_NVMAX || !cfg->zlev || !(cfg->dx > 0) || !(cfg->dy > 0) ||
        !(cfg->sigma_h > 0) || !(cfg->sigma_v > 0))
        return fail(LEKF_ERR_SIZE);
    if (lekf_device_count() < 1) return fail(LEKF_ERR_NODEV);
    lekf_handle* h = new lekf_handle();
    h->kind = 2;
    h->r = *cfg;
    h->zlev.assign(cfg->zlev, cfg->zlev + cfg->nlev);
    h->r.zlev = h->zlev.data();
    h->ne = cfg->ne;
    h->nij = cfg->nx * cfg->ny;
    h->nlev = cfg->nlev;
    h->nv3d = cfg->nv3d;
    h->nv2d = 0;
    h->prec = cfg->prec;
    h->infl_mode = cfg->infl_mode;
    h->infl_value = cfg->infl_value;
    if (cfg->list_budget > 0) h->list_budget = cfg->list_budget;
    // search rectangle half-widths in grid units (scale/letkf/letkf_obs.f90:122-123)
    const double dist_zero_fac = 3.651483717;
    h->dlon_zero = cfg->sigma_h * dist_zero_fac / cfg->dx;
    h->dlat_zero = cfg->sigma_h * dist_zero_fac / cfg->dy;
    // the selection kernel keeps the rectangle's row segments in LDS
    if (std::min<double>(cfg->ny, 2.0 * h->dlat_zero + 3.0) > LEKF_REG_SEG) {
        delete h;
        return fail(LEKF_ERR_SIZE);
    }
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(LEKF_ERR_HIP);
    }
    int rc = upload(h->d_zlev, cfg->zlev, sizeof(double) * cfg->nlev, h->st);
    if (!rc) rc = alloc_state(h);
    if (rc) {
        free_all(h);
        (void)hipStreamDestroy(h->st);
        delete h;
        return fail(rc);
    }
    if (status) *status = 0;
    return h;
}

// The regional obs index from the obs in the handle's r_* buffers (ri, rj, z,
// err, dep, bucket, qc if has_qc (obs kept where it equals qc_keep),
// perturbations of member m of obs n at r_hxf[n * h_so + m * h_sm]): SCALE's bin sort by grid cell
// (scale/letkf/letkf_obs.f90:733-830), the per-point list capacities and the
// chunking of the points by the list budget.
static int regional_obs_index(lekf_handle* h, int nobs, bool has_qc, int qc_keep, long h_so,
                              long h_sm)
{
    hipStream_t st = h->st;
    const int ne = h->ne, nx = h->r.nx, ny = h->r.ny;
    const size_t n1 = (size_t)(nobs > 0 ? nobs : 1);
    long n2 = 2048;
    while (n2 < nobs) n2 <<= 1;
    const int ncell = nx * ny;
    HIPCHK(h->keys.ensure(sizeof(unsigned long long) * n2));
    HIPCHK(h->cell_cnt.ensure(sizeof(int) * ncell));
    HIPCHK(h->cellcum.ensure(sizeof(int) * ncell));
    HIPCHK(h->perm.ensure(sizeof(int) * n1));
    HIPCHK(h->o_ri.ensure(sizeof(double) * n1));
    HIPCHK(h->o_rj.ensure(sizeof(double) * n1));
    HIPCHK(h->o_z.ensure(sizeof(double) * n1));
    HIPCHK(h->o_err.ensure(sizeof(double) * n1));
    HIPCHK(h->o_rdiag.ensure(sizeof(double) * n1));
    HIPCHK(h->o_dep.ensure(sizeof(double) * n1));
    HIPCHK(h->o_bkt.ensure(sizeof(int) * n1));
    HIPCHK(h->o_hdxf.ensure(sizeof(double) * n1 * ne));
    if (h->prec != LEKF_PREC_F64) HIPCHK(h->o_hdxf32.ensure(sizeof(float) * n1 * ne));
    HIPCHK(hipMemsetAsync(h->cell_cnt.p, 0, sizeof(int) * ncell, st));
    HIPCHK(hipMemsetAsync(h->keys.p, 0xff, sizeof(unsigned long long) * n2, st));
    RegObsArgs A;
    memset(&A, 0, sizeof(A));
    A.nobs = nobs;
    A.ne = ne;
    A.nx = nx;
    A.ny = ny;
    A.ri = h->r_lon.as<double>();
    A.rj = h->r_lat.as<double>();
    A.z = h->r_z.as<double>();
    A.err = h->r_err.as<double>();
    A.dep = h->r_dep.as<double>();
    A.bkt = h->r_bkt.as<int>();
    A.qc = has_qc ? h->r_qc.as<int>() : nullptr;
    A.qc_keep = qc_keep;
    A.hdxf = h->r_hxf.as<double>();
    A.h_so = h_so;
    A.h_sm = h_sm;
    A.keys = h->keys.as<unsigned long long>();
    A.cell_cnt = h->cell_cnt.as<int>();
    A.perm = h->perm.as<int>();
    A.o_ri = h->o_ri.as<double>();
    A.o_rj = h->o_rj.as<double>();
    A.o_z = h->o_z.as<double>();
    A.o_err = h->o_err.as<double>();
    A.o_rdiag = h->o_rdiag.as<double>();
    A.o_dep = h->o_dep.as<double>();
    A.o_hdxf = h->o_hdxf.as<double>();
    A.o_hdxf32 = h->prec != LEKF_PREC_F64 ? h->o_hdxf32.as<float>() : nullptr;
    A.o_bkt = h->o_bkt.as<int>();
    HIPCHK(lekf_reg_keys_launch(A, st));
    HIPCHK(lekf_sort_keys_launch(A.keys, n2, st));
    std::vector<int> cc(ncell);
    HIPCHK(hipMemcpyAsync(cc.data(), h->cell_cnt.p, sizeof(int) * ncell, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    long nbin = 0;
    for (int c : cc) nbin += c;
    h->nbinned = (int)nbin;
    h->nobs = nobs;
    HIPCHK(lekf_reg_scatter_launch(A, h->nbinned, ncell, h->cellcum.as<int>(), st));
    if (lekf_reg_select_fused(h->r.nobs_max)) {
        const size_t nb1 = (size_t)(h->nbinned > 0 ? h->nbinned : 1);
        HIPCHK(h->b_cnt.ensure(sizeof(int) * (size_t)h->nbkt * ncell));
        HIPCHK(h->b_cc.ensure(sizeof(int) * (size_t)h->nbkt * ncell));
        HIPCHK(h->b_rec.ensure(4 * sizeof(double) * nb1));
        HIPCHK(h->b_o.ensure(sizeof(int) * nb1));
        HIPCHK(h->wc_so.ensure(sizeof(int) * LEKF_CAP_LDS * (size_t)lekf_reg_wave_slots_max()));
        HIPCHK(h->b_err2min.ensure(sizeof(unsigned long long) * LEKF_REG_NB));
        HIPCHK(hipMemsetAsync(h->b_err2min.p, 0xff, sizeof(unsigned long long) * LEKF_REG_NB, st));
        HIPCHK(lekf_reg_buckets_launch(A, h->nbinned, h->cellcum.as<int>(), h->nbkt,
                                       h->b_cnt.as<int>(), h->b_cc.as<int>(), h->b_rec.as<double>(),
                                       h->b_o.as<int>(), h->b_err2min.as<unsigned long long>(), st));
    }
    // list capacity per point -> chunks of points whose lists fit the budget
    const long npt = (long)h->nij * h->nlev;
    HIPCHK(h->cap_n.ensure(sizeof(int) * npt));
    HIPCHK(h->cap_off.ensure(sizeof(long) * (npt + 1)));
    RegionalArgs R = make_reg_args(h);
    HIPCHK(h->maxcand.ensure(sizeof(int)));
    HIPCHK(hipMemsetAsync(h->maxcand.p, 0, sizeof(int), st));
    HIPCHK(lekf_reg_capacity_launch(R, h->cap_n.as<int>(), h->cap_off.as<long>(),
                                    h->maxcand.as<int>(), st));
    int maxcand = 0;
    HIPCHK(hipMemcpyAsync(&maxcand, h->maxcand.p, sizeof(int), hipMemcpyDeviceToHost, st));
    std::vector<long> off(npt + 1);
    HIPCHK(hipMemcpyAsync(off.data(), h->cap_off.p, sizeof(long) * (npt + 1), hipMemcpyDeviceToHost,
                          st));
    HIPCHK(hipStreamSynchronize(st));
    h->chunk_pts.assign(1, 0);
    h->chunk_off.assign(1, 0);
    h->held_chunk = -1;
    long maxlist = 1;
    for (long p = 0; p < npt;) {
        long q = p + 1;
        while (q < npt && off[q + 1] - off[p] <= h->list_budget) ++q;
        h->chunk_pts.push_back(q);
        h->chunk_off.push_back(off[q]);
        maxlist = std::max(maxlist, off[q] - off[p]);
        p = q;
    }
    HIPCHK(h->loc_idx.ensure(sizeof(int) * maxlist));
    HIPCHK(h->loc_rho.ensure(sizeof(double) * maxlist));
    h->loc_cap = maxlist;
    // capped selection: per workgroup slot one stream per bucket, each as
    // long as the largest rectangle's candidate count
    h->sel_slots = lekf_reg_select_slots_max();
    h->slot_cap = h->r.nobs_max > 0 && !lekf_reg_select_fused(h->r.nobs_max)
                      ? std::max(maxcand, 1) : 0;
    if (h->slot_cap > 0) {
        const size_t ne_ = (size_t)h->sel_slots * (size_t)h->nbkt * (size_t)h->slot_cap;
        HIPCHK(h->sel_key.ensure(sizeof(unsigned long long) * ne_));
        HIPCHK(h->sel_ob.ensure(sizeof(int) * ne_));
        HIPCHK(h->sel_rho.ensure(sizeof(double) * ne_));
        HIPCHK(ensure_cap_sets(h, h->r.nobs_max));
    }
    h->have_obs = true;
    return upload_jobs(h);
}


int lekf_set_obs_regional(lekf_handle* h, int nobs, const double* ri, const double* rj,
                          const double* z, const double* err, const double* dep, const int* bucket,
                          const int* qc, const double* hdxf)
{
    if (!h || h->kind != 2 || nobs < 0) return LEKF_ERR_ARG;
    if (h) graph_drop(h);
    if (nobs > 0 && (!ri || !rj || !z || !err || !dep || !bucket || !hdxf)) return LEKF_ERR_ARG;
    hipStream_t st = h->st;
    const int ne = h->ne;
    int nb = 1;
    for (int n = 0; n < nobs; ++n) {
        if (bucket[n] < 0 || bucket[n] >= LEKF_REG_NB) return LEKF_ERR_ARG;
        nb = std::max(nb, bucket[n] + 1);
    }
    h->nbkt = nb;
    int rc;
    if ((rc = upload(h->r_lon, ri, sizeof(double) * nobs, st)) ||
        (rc = upload(h->r_lat, rj, sizeof(double) * nobs, st)) ||
        (rc = upload(h->r_z, z, sizeof(double) * nobs, st)) ||
        (rc = upload(h->r_err, err, sizeof(double) * nobs, st)) ||
        (rc = upload(h->r_dep, dep, sizeof(double) * nobs, st)) ||
        (rc = upload(h->r_bkt, bucket, sizeof(int) * nobs, st)) ||
        (rc = upload(h->r_hxf, hdxf, sizeof(double) * nobs * (size_t)ne, st)))
        return rc;
    if (qc && (rc = upload(h->r_qc, qc, sizeof(int) * nobs, st))) return rc;
    return regional_obs_index(h, nobs, qc != nullptr, 1, 1, nobs);
}

// Radar obs straight from the resident ensemble: the radar operator, SCALE's
// departure and QC step and the regional obs index, with H(x), the
// perturbations and the departures kept on the device between them.
int lekf_set_obs_radar(lekf_handle* h, const lekf_radar_cfg* rcfg, const lekf_scale_qc_cfg* qcfg,
                       int nobs, const double* elm, const double* lon, const double* lat,
                       const double* z, const double* ri, const double* rj, const double* radar_lon,
                       const double* radar_lat, const double* radar_z, const double* dat,
                       const double* err, const int* bucket, int* qc_out)
{
    if (!h || !rcfg || !qcfg || h->kind != 2 || nobs < 0 || !h->have_ens || !h->d_hgt.p ||
        h->nlev < 2)
        return LEKF_ERR_ARG;
    if (h) graph_drop(h);
    if (rcfg->method < 1 || rcfg->method > 3) return LEKF_ERR_ARG;
    for (int q = 0; q < 11; ++q)
        if (rcfg->var[q] < 0 || rcfg->var[q] >= h->nv3d) return LEKF_ERR_ARG;
    const double* in[11] = {elm, lon, lat, z, ri, rj, radar_lon, radar_lat, radar_z, dat, err};
    if (nobs > 0) {
        for (auto* p : in)
            if (!p) return LEKF_ERR_ARG;
        if (!bucket) return LEKF_ERR_ARG;
    }
    int nb = 1;
    for (int n = 0; n < nobs; ++n) {
        if (bucket[n] < 0 || bucket[n] >= LEKF_REG_NB) return LEKF_ERR_ARG;
        nb = std::max(nb, bucket[n] + 1);
    }
    h->nbkt = nb;
    hipStream_t st = h->st;
    const size_t n = (size_t)nobs, ne = (size_t)h->ne, n1 = n > 0 ? n : 1;
    DevBuf rec, dqm;
    struct Release {  // the call's buffers go on every return path
        DevBuf* b[2];
        ~Release()
        {
            for (DevBuf* x : b) x->release();
        }
    } rel{{&rec, &dqm}};
    HIPCHK(rec.ensure(sizeof(double) * n1 * 12));
    double* r = rec.as<double>();
    for (int f = 0; f < 11 && n > 0; ++f)
        HIPCHK(hipMemcpyAsync(r + n * f, in[f], sizeof(double) * n, hipMemcpyHostToDevice, st));
    HIPCHK(dqm.ensure(sizeof(int) * n1 * ne));
    HIPCHK(h->r_hxf.ensure(sizeof(double) * n1 * ne));
    int rc;
    if ((rc = upload(h->r_lon, ri, sizeof(double) * n, st)) ||
        (rc = upload(h->r_lat, rj, sizeof(double) * n, st)) ||
        (rc = upload(h->r_z, z, sizeof(double) * n, st)) ||
        (rc = upload(h->r_bkt, bucket, sizeof(int) * n, st)))
        return rc;
    HIPCHK(h->r_err.ensure(sizeof(double) * n1));
    HIPCHK(h->r_dep.ensure(sizeof(double) * n1));
    HIPCHK(h->r_qc.ensure(sizeof(int) * n1));
    RadarArgs A{};
    A.nobs = nobs;
    A.nx = h->r.nx;
    A.ny = h->r.ny;
    A.nlev = h->nlev;
    A.ne = h->ne;
    A.nv3d = h->nv3d;
    for (int q = 0; q < 11; ++q) A.var[q] = rcfg->var[q];
    A.method = rcfg->method;
    A.use_wt = rcfg->use_terminal_velocity ? 1 : 0;
    A.min_ref_dbz = rcfg->min_ref_dbz;
    A.low_ref_shift = rcfg->low_ref_shift;
    A.zmax = rcfg->zmax;
    A.gam[0] = std::tgamma(4.8);
    A.gam[1] = std::tgamma(4.25);
    A.gam[2] = std::tgamma(4.5);
    A.elm = r;
    A.lon = r + n;
    A.lat = r + 2 * n;
    A.z = r + 3 * n;
    A.ri = r + 4 * n;
    A.rj = r + 5 * n;
    A.rlon = r + 6 * n;
    A.rlat = r + 7 * n;
    A.rz = r + 8 * n;
    A.x3 = h->x3b.as<double>();
    A.hgt = h->d_hgt.as<double>();
    A.hxf = h->r_hxf.as<double>();
    A.qc = dqm.as<int>();
    A.rk = r + 11 * n;
    HIPCHK(lekf_radar_launch(A, st));
    auto factor = [&](double f) { return f < 0.0 ? qcfg->gross_error : f; };
    ScaleQcArgs Q{};
    Q.nobs = nobs;
    Q.ne = h->ne;
    Q.use_ref = qcfg->use_radar_ref != 0;
    Q.use_vr = qcfg->use_radar_vr != 0;
    Q.use_obserr_ref = qcfg->use_obserr_radar_ref != 0;
    Q.use_obserr_vr = qcfg->use_obserr_radar_vr != 0;
    Q.min_member = qcfg->min_radar_ref_member;
    Q.min_member_obsref = qcfg->min_radar_ref_member_obsref;
    Q.thres = std::max(qcfg->radar_ref_thres_dbz, qcfg->min_radar_ref_dbz);
    Q.min_ref_dbz = qcfg->min_radar_ref_dbz;
    Q.low_ref_shift = qcfg->low_ref_shift;
    Q.min_ref = std::pow(10.0, qcfg->min_radar_ref_dbz / 10.0);
    Q.obserr_ref = qcfg->obserr_radar_ref;
    Q.obserr_vr = qcfg->obserr_radar_vr;
    Q.gross = qcfg->gross_error;
    Q.gross_rain = factor(qcfg->gross_error_rain);
    Q.gross_ref = factor(qcfg->gross_error_radar_ref);
    Q.gross_vr = factor(qcfg->gross_error_radar_vr);
    Q.gross_prh = factor(qcfg->gross_error_radar_prh);
    Q.gross_tcx = factor(qcfg->gross_error_tcx);
    Q.gross_tcy = factor(qcfg->gross_error_tcy);
    Q.gross_tcp = factor(qcfg->gross_error_tcp);
    Q.elm = r;
    Q.dat_in = r + 9 * n;
    Q.err_in = r + 10 * n;
    Q.so = h->ne;  // the operator's rows: member m of obs n at n * ne + m
    Q.sm = 1;
    Q.qc_mem = dqm.as<int>();
    Q.hx = h->r_hxf.as<double>();
    Q.dat = r + 11 * n;  // the dBZ values are not kept: rk's slot is free again
    Q.err = h->r_err.as<double>();
    Q.dep = h->r_dep.as<double>();
    Q.qc = h->r_qc.as<int>();
    HIPCHK(lekf_scale_qc_launch(Q, st));
    if (qc_out && n > 0)
        HIPCHK(hipMemcpyAsync(qc_out, h->r_qc.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    return regional_obs_index(h, nobs, true, LEKF_IQC_GOOD, (long)ne, 1);
}

int lekf_set_stream(lekf_handle* h, void* stream)
{
    if (!h) return LEKF_ERR_ARG;
    HIPCHK(hipStreamSynchronize(h->st));
    if (h->own_stream) HIPCHK(hipStreamDestroy(h->st));
    h->own_stream = false;
    h->st = (hipStream_t)stream;
    return LEKF_OK;
}

namespace {
struct BatchBufs {
    DevBuf jobs[3], ws;
};
std::mutex g_batch_mu;
std::map<hipStream_t, BatchBufs> g_batch_bufs;  // per stream: batches on other streams overlap

// Enqueue the selection of every handle and the grouped solves on st.
int run_batch_on(lekf_handle** hs, int n, hipStream_t st)
{
    for (int i = 0; i < n; ++i)
        if (!hs[i] || !hs[i]->have_ens || !hs[i]->have_obs) return LEKF_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_batch_mu);
    BatchBufs& B = g_batch_bufs[st];
    std::vector<lekf_handle*> batched;
    for (int i = 0; i < n; ++i) {
        lekf_handle* h = hs[i];
        if (h->st != st) {
            HIPCHK(hipStreamSynchronize(h->st));
            if (h->own_stream) HIPCHK(hipStreamDestroy(h->st));
            h->own_stream = false;
            h->st = st;
        }
        if (h->kind == 2 && h->chunk_pts.size() != 2) {
            // lists built chunk by chunk: select and solve per chunk, on its own
            const int rc = lekf_das_run(h);
            if (rc) return rc;
            continue;
        }
        batched.push_back(h);
        const int rc = run_select(h);
        if (rc) return rc;
    }
    // one grouped launch set per arithmetic mode
    for (int mode = 0; mode < 3; ++mode) {
        std::vector<lekf_job> J;
        for (lekf_handle* h : batched)
            if (mode_of_prec(h->prec) == mode) J.push_back(h->job);
        if (J.empty()) continue;
        const int rc = launch_grouped(J, mode, B.jobs[mode], B.ws, st);
        if (rc) return rc;
    }
    for (lekf_handle* h : batched) {
        const int rc = run_relax(h);
        if (rc) return rc;
    }
    return LEKF_OK;
}
}  // namespace

int lekf_das_run_batch(lekf_handle** hs, int n, void* stream)
{
    if (!hs || n <= 0) return LEKF_ERR_ARG;
    return run_batch_on(hs, n, (hipStream_t)stream);
}

int lekf_das_run_many(lekf_handle** hs, int n)
{
    if (!hs || n <= 0 || !hs[0]) return LEKF_ERR_ARG;
    // selection and solves on the leading handle's stream; the others keep theirs
    hipStream_t st = hs[0]->st;
    std::vector<hipStream_t> keep(n);
    std::vector<bool> own(n);
    for (int i = 0; i < n; ++i) {
        if (!hs[i]) return LEKF_ERR_ARG;
        keep[i] = hs[i]->st;
        own[i] = hs[i]->own_stream;
        if (i > 0 && hs[i]->st != st) {
            HIPCHK(hipStreamSynchronize(hs[i]->st));
            hs[i]->own_stream = false;  // borrowed for the batch, restored below
            hs[i]->st = st;
        }
    }
    const int rc = run_batch_on(hs, n, st);
    const hipError_t e = hipStreamSynchronize(st);
    for (int i = 0; i < n; ++i) {
        hs[i]->st = keep[i];
        hs[i]->own_stream = own[i];
    }
    if (rc) return rc;
    HIPCHK(e);
    return LEKF_OK;
}

int lekf_sync(lekf_handle* h)
{
    if (!h) return LEKF_ERR_ARG;
    HIPCHK(hipStreamSynchronize(h->st));
    return LEKF_OK;
}

int lekf_get_analysis(lekf_handle* h, double* anal3d, double* anal2d)
{
    if (!h || !anal3d) return LEKF_ERR_ARG;
    hipStream_t st = h->st;
    const size_t n3 = (size_t)h->nij * h->nlev * h->ne * h->nv3d;
    const size_t n2 = (size_t)h->nij * h->ne * h->nv2d;
    HIPCHK(h->stage.ensure(sizeof(double) * std::max(n3, n2)));
    hipLaunchKernelGGL(ens3_from_dev, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st,
                       h->x3a.as<double>(), h->stage.as<double>(), h->nij, h->nlev, h->ne, h->nv3d);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(anal3d, h->stage.p, sizeof(double) * n3, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (n2 && anal2d) {
        hipLaunchKernelGGL(ens3_from_dev, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st,
                           h->x2a.as<double>(), h->stage.as<double>(), h->nij, 1, h->ne, h->nv2d);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(anal2d, h->stage.p, sizeof(double) * n2, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    return LEKF_OK;
}

int lekf_get_infl(lekf_handle* h, double* infl3d, double* infl2d)
{
    if (!h || !infl3d) return LEKF_ERR_ARG;
    const size_t npt = (size_t)h->nij * h->nlev;
    HIPCHK(hipMemcpyAsync(infl3d, h->infl_out.p, sizeof(double) * npt, hipMemcpyDeviceToHost, h->st));
    if (infl2d && h->nv2d)
        HIPCHK(hipMemcpyAsync(infl2d, h->infl2.p, sizeof(double) * h->nij * h->nv2d,
                              hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return LEKF_OK;
}

int lekf_get_nobsl(lekf_handle* h, int* nobsl)
{
    if (!h || !nobsl) return LEKF_ERR_ARG;
    const size_t npt = (size_t)h->nij * h->nlev;
    HIPCHK(hipMemcpyAsync(nobsl, h->job.loc_cnt, sizeof(int) * npt, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return LEKF_OK;
}

int lekf_get_local(lekf_handle* h, long b, int cap, int* idx, double* rho)
{
    if (!h || b < 0 || b >= (long)h->nij * h->nlev) return -LEKF_ERR_ARG;
    if (h->kind == 2 && h->have_obs) {
        // the list buffer holds one chunk of points: select the chunk of b again if needed
        int c = 0;
        while (h->chunk_pts[c + 1] <= b) ++c;
        if (c != h->held_chunk) {
            if (hipStreamSynchronize(h->st) != hipSuccess) return -LEKF_ERR_HIP;
            const int rc = reg_select_chunk(h, c);
            if (rc) return -rc;
        }
        if (hipStreamSynchronize(h->st) != hipSuccess) return -LEKF_ERR_HIP;
    }
    if (!h->have_obs) return -LEKF_ERR_ARG;
    if (hipStreamSynchronize(h->st) != hipSuccess) return -LEKF_ERR_HIP;
    // the lists the solve reads (capped ones for a global grid with nobs_max)
    const lekf_job& J = h->job;
    long off = 0;
    int cnt = 0;
    if (hipMemcpy(&off, J.loc_off + b, sizeof(long), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&cnt, J.loc_cnt + b, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return -LEKF_ERR_HIP;
    const int m = cnt < cap ? cnt : cap;
    if (m > 0) {
        if (hipMemcpy(idx, J.loc_idx + off, sizeof(int) * m, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(rho, J.loc_rho + off, sizeof(double) * m, hipMemcpyDeviceToHost) != hipSuccess)
            return -LEKF_ERR_HIP;
    }
    return cnt;
}

int lekf_get_obs_index(lekf_handle* h, int* perm, int* nobsgrd, int* nbinned)
{
    if (!h || (h->kind != 0 && h->kind != 2)) return LEKF_ERR_ARG;
    HIPCHK(hipStreamSynchronize(h->st));
    if (nbinned) *nbinned = h->nbinned;
    if (perm && h->nbinned)
        HIPCHK(hipMemcpy(perm, h->perm.p, sizeof(int) * h->nbinned, hipMemcpyDeviceToHost));
    if (nobsgrd && h->kind == 0)
        HIPCHK(hipMemcpy(nobsgrd, h->nobsgrd.p, sizeof(int) * h->g.nlon * h->g.nlat,
                         hipMemcpyDeviceToHost));
    if (nobsgrd && h->kind == 2 && h->have_obs)
        HIPCHK(hipMemcpy(nobsgrd, h->cellcum.p, sizeof(int) * h->nij, hipMemcpyDeviceToHost));
    return LEKF_OK;
}

int lekf_get_flags(lekf_handle* h, long* n_rankdef, long* n_noconv)
{
    if (!h) return LEKF_ERR_ARG;
    const size_t npt = (size_t)h->nij * h->nlev;
    std::vector<int> s(npt);
    HIPCHK(hipMemcpyAsync(s.data(), h->status.p, sizeof(int) * npt, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    long a = 0, b = 0;
    for (int v : s) {
        a += (v & LEKF_PT_RANKDEF) ? 1 : 0;
        b += (v & LEKF_PT_NOCONV) ? 1 : 0;
    }
    if (n_rankdef) *n_rankdef = a;
    if (n_noconv) *n_noconv = b;
    return LEKF_OK;
}

int lekf_get_fp64_count(lekf_handle* h, long* n_fp64)
{
    if (!h || !n_fp64) return LEKF_ERR_ARG;
    const size_t npt = (size_t)h->nij * h->nlev;
    std::vector<int> s(npt);
    HIPCHK(hipMemcpyAsync(s.data(), h->status.p, sizeof(int) * npt, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    long a = 0;
    for (int v : s) a += (v & LEKF_PT_FP64) ? 1 : 0;
    *n_fp64 = a;
    return LEKF_OK;
}

int lekf_cycle_swap(lekf_handle* h)
{
    if (!h) return LEKF_ERR_ARG;
    HIPCHK(hipStreamSynchronize(h->st));
    graph_drop(h);
    std::swap(h->x3b, h->x3a);
    std::swap(h->x2b, h->x2a);
    std::swap(h->infl_in, h->infl_out);
    if (h->have_obs) return upload_jobs(h);
    return LEKF_OK;
}

int lekf_l96_forecast(lekf_handle* h, int kt, double dt, double force)
{
    if (!h || h->kind != 1 || kt < 0 || !h->have_ens) return LEKF_ERR_ARG;
    HIPCHK(hipStreamSynchronize(h->st));
    graph_drop(h);
    HIPCHK(lekf_l96_rk4_launch(h->x3a.as<double>(), h->x3b.as<double>(), h->nij, h->ne, kt, dt,
                               force, h->st));
    std::swap(h->infl_in, h->infl_out);
    if (h->have_obs) return upload_jobs(h);
    return LEKF_OK;
}

int lekf_device_state(lekf_handle* h, double** x3b, double** x3a, double** x2b, double** x2a)
{
    if (!h) return LEKF_ERR_ARG;
    if (x3b) *x3b = h->x3b.as<double>();
    if (x3a) *x3a = h->x3a.as<double>();
    if (x2b) *x2b = h->x2b.as<double>();
    if (x2a) *x2a = h->x2a.as<double>();
    return LEKF_OK;
}

}  // extern "C"
