// capi.cpp -- the C ABI of include/openpiv_amd.h over the HIP kernels.
//
// Host-side replacement of the per-window loop of
// examples/process/main.cpp:137-302: the grid is generated once per plan
// (grid.impl.h:10-61), and one kernel launch processes every window of a
// frame pair.  Error texts that mirror reference exceptions use the
// reference's wording.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "adjoint.hpp"
#include "background.hpp"
#include "corr_direct.hpp"
#include "corr_fft.hpp"
#include "ctx.hpp"
#include "openpiv_amd.h"
#include "plan_check.hpp"
#include "transform.hpp"

using namespace opiv::detail;

namespace {

// Scratch of opiv_vjp_pair_async_noalloc: the VJP's per-window tap weights
// (32 B each) for callers that cannot allocate.  One copy per GPU, held until
// the process ends; every context on a GPU shares it.  Host code never uses
// this name as an address: the device address of each GPU's copy is looked
// up once per device (scratch_address) and kept.
static __device__ __align__(256) opiv::AdjCoef scratch[OPIV_NOALLOC_MAX_WINDOWS];

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;      // device -> address of its copy of scratch

// the address of the current device's copy (resolved on the first call per
// device, with that device current)
int scratch_address(int device, void** out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_ptr[(size_t)device]) {
        void* ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[(size_t)device] = ptr;
    }
    *out = g_scratch_ptr[(size_t)device];
    return OPIV_OK;
}

} // namespace

extern "C" {

const char* opiv_last_error(void) { return g_last_error.c_str(); }

const char* opiv_version(void) { return "openpiv-amd 0.1.0 (gfx950)"; }

int opiv_ctx_create(int device, opiv_ctx** out)
{
    if (!out) return fail(OPIV_E_INVALID, "opiv_ctx_create: null output pointer");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return fail(OPIV_E_NO_DEVICE, "no HIP device available");
    }
    if (device < 0 || device >= count)
        return fail(OPIV_E_INVALID, "device %d out of range (%d devices)", device, count);
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(OPIV_E_NO_DEVICE, "device %d is %s; this library is built for gfx950 only", device, prop.gcnArchName);
    HIP_TRY(hipSetDevice(device));
    void* scratch_ptr = nullptr;
    if (int rc = scratch_address(device, &scratch_ptr)) return rc;
    opiv_ctx* c = new opiv_ctx;
    c->device = device;
    c->scratch_coef = scratch_ptr;
    c->cu_count = prop.multiProcessorCount;
    c->arch = prop.gcnArchName;
    c->arch = c->arch.substr(0, c->arch.find(':'));   // "gfx950:sramecc+:xnack-" -> "gfx950"
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "hipStreamCreateWithFlags");
    }
    *out = c;
    return OPIV_OK;
}

int opiv_ctx_destroy(opiv_ctx* ctx)
{
    if (!ctx) return OPIV_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto* b : { &ctx->stage_a, &ctx->stage_b, &ctx->stage_out, &ctx->stage_fwd, &ctx->stage_cot,
                     &ctx->stage_ga, &ctx->stage_gb, &ctx->stage_var, &ctx->adj_coef, &ctx->bg_frames,
                     &ctx->bg_out, &ctx->bg_mean, &ctx->bg_table, &ctx->bg_range })
        b->release();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return OPIV_OK;
}

int opiv_ctx_synchronize(opiv_ctx* ctx)
{
    if (!ctx) return fail(OPIV_E_INVALID, "opiv_ctx_synchronize: null context");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return OPIV_OK;
}

void* opiv_ctx_stream(opiv_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int opiv_plan_create(opiv_ctx* ctx, const opiv_plan_desc* desc, opiv_plan** out)
{
    if (!ctx || !desc || !out) return fail(OPIV_E_INVALID, "opiv_plan_create: null argument");
    *out = nullptr;
    PlanGeom g;
    if (int rc = check_plan_desc(desc, &g)) return rc;
    opiv_plan* p = new opiv_plan;
    p->ctx = ctx;
    p->desc = *desc;
    p->desc.pitch = g.pitch;
    p->desc.variant = g.variant;
    p->off_x = g.off_x;
    p->off_y = g.off_y;
    p->nx = g.nx;
    p->ny = g.ny;
    p->x_start = g.x_start;
    p->y_start = g.y_start;
    *out = p;
    return OPIV_OK;
}

int opiv_plan_destroy(opiv_plan* plan)
{
    delete plan;
    return OPIV_OK;
}

int opiv_plan_window_count(const opiv_plan* plan, size_t* count)
{
    if (!plan || !count) return fail(OPIV_E_INVALID, "opiv_plan_window_count: null argument");
    *count = (size_t)plan->nx * plan->ny;
    return OPIV_OK;
}

int opiv_plan_grid(const opiv_plan* plan, int32_t* xy)
{
    if (!plan || !xy) return fail(OPIV_E_INVALID, "opiv_plan_grid: null argument");
    uint32_t nx, ny;
    return opiv_grid(plan->desc.img_w, plan->desc.img_h, plan->desc.win, plan->desc.win,
                     plan->off_x, plan->off_y, &nx, &ny, xy);
}

// Device frames must be aligned to their pixel size: 8- and 16-bit pixels
// are read by typed buffer loads, which take the address of a whole element.
static int check_frames(const opiv_plan* plan, const void* da, const void* db, const char* who)
{
    const size_t es = dtype_size(plan->desc.dtype);
    if (((uintptr_t)da | (uintptr_t)db) % es != 0)
        return fail(OPIV_E_INVALID, "%s: frame pointers must be aligned to their %zu-byte pixels", who, es);
    return OPIV_OK;
}

static opiv::CorrParams plan_params(const opiv_plan* plan, const void* a, const void* b, opiv_window_result* out)
{
    opiv::CorrParams p{};
    p.a = a;
    p.b = b;
    p.pitch = plan->desc.pitch;
    p.x_start = plan->x_start;
    p.y_start = plan->y_start;
    p.step_x = (int32_t)plan->off_x;
    p.step_y = (int32_t)plan->off_y;
    p.nx = (int32_t)plan->nx;
    p.nx_mul = opiv::fastdiv_mul(plan->nx);
    p.first = 0;
    p.n_windows = (int32_t)(plan->nx * plan->ny);
    p.img_h = plan->desc.img_h;
    p.limit_search = plan->desc.limit_search ? 1 : 0;
    p.real_compat = plan->desc.variant == OPIV_REAL_COMPAT ? 1 : 0;
    p.tie_tol = kTieTol;
    p.out = out;
    p.planes = nullptr;
    return p;
}

// the plan's forward kernel
static hipError_t launch_plan(const opiv_plan* plan, const opiv::CorrParams& p, hipStream_t stream)
{
    if (plan->desc.variant == OPIV_DIRECT_BF16_SPLIT)
        return opiv::launch_corr_direct(p, (int)plan->desc.win, plan->desc.dtype, false, stream);
    return opiv::launch_corr_fft(p, (int)plan->desc.win, plan->desc.dtype, false, stream);
}

int opiv_run_pair_async(opiv_plan* plan, const void* d_a, const void* d_b, opiv_window_result* d_out, size_t n)
{
    if (!plan || !d_a || !d_b || !d_out) return fail(OPIV_E_INVALID, "opiv_run_pair_async: null argument");
    const size_t count = (size_t)plan->nx * plan->ny;
    if (n != count) return fail(OPIV_E_SIZE, "result count %zu differs from the grid's %zu windows", n, count);
    if (int rc = check_frames(plan, d_a, d_b, "opiv_run_pair_async")) return rc;
    if (!is_device_ptr(d_a) || !is_device_ptr(d_b) || !is_device_ptr(d_out))
        return fail(OPIV_E_INVALID, "opiv_run_pair_async: frames and results must be device memory");
    HIP_TRY(hipSetDevice(plan->ctx->device));
    opiv::CorrParams p = plan_params(plan, d_a, d_b, d_out);
    HIP_TRY(launch_plan(plan, p, plan->ctx->stream));
    return OPIV_OK;
}

int opiv_run_pair(opiv_plan* plan, const void* a, const void* b, opiv_window_result* out, size_t n)
{
    if (!plan || !a || !b || !out) return fail(OPIV_E_INVALID, "opiv_run_pair: null argument");
    opiv_ctx* ctx = plan->ctx;
    const size_t count = (size_t)plan->nx * plan->ny;
    if (n != count) return fail(OPIV_E_SIZE, "result count %zu differs from the grid's %zu windows", n, count);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t frame_bytes = plan->desc.pitch * plan->desc.img_h * dtype_size(plan->desc.dtype);
    const void* da = a;
    const void* db = b;
    if (!is_device_ptr(a)) {
        HIP_TRY(ctx->stage_a.ensure(frame_bytes));
        HIP_TRY(hipMemcpyAsync(ctx->stage_a.p, a, frame_bytes, hipMemcpyHostToDevice, ctx->stream));
        da = ctx->stage_a.p;
    }
    if (!is_device_ptr(b)) {
        HIP_TRY(ctx->stage_b.ensure(frame_bytes));
        HIP_TRY(hipMemcpyAsync(ctx->stage_b.p, b, frame_bytes, hipMemcpyHostToDevice, ctx->stream));
        db = ctx->stage_b.p;
    }
    if (int rc = check_frames(plan, da, db, "opiv_run_pair")) return rc;
    const bool out_dev = is_device_ptr(out);
    opiv_window_result* dout = out;
    if (!out_dev) {
        HIP_TRY(ctx->stage_out.ensure(count * sizeof(opiv_window_result)));
        dout = (opiv_window_result*)ctx->stage_out.p;
    }
    opiv::CorrParams p = plan_params(plan, da, db, dout);
    HIP_TRY(launch_plan(plan, p, ctx->stream));
    if (!out_dev)
        HIP_TRY(hipMemcpyAsync(out, dout, count * sizeof(opiv_window_result), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return OPIV_OK;
}

static int correlate_stacked(opiv_ctx* ctx, uint32_t win, size_t n, int variant,
                             const double* a, const double* b, double* planes)
{
    if (!ctx || !a || !b || !planes) return fail(OPIV_E_INVALID, "correlate: null argument");
    if (int rc = check_window(win)) return rc;
    if (variant < OPIV_CIRCULAR_FFT_FP32 || variant > OPIV_AUTO)
        return fail(OPIV_E_UNSUPPORTED, "unknown variant %d", variant);
    if (variant == OPIV_AUTO) variant = auto_variant((int)win, OPIV_F64);
    if (variant == OPIV_DIRECT_BF16_SPLIT && !opiv::direct_supported((int)win, OPIV_F64))
        return fail(OPIV_E_UNSUPPORTED, "the direct path does not implement %ux%u windows of FP64 values", win, win);
    if (n == 0) return OPIV_OK;
    if (n > 0x7fffffffull / win) return fail(OPIV_E_UNSUPPORTED, "too many windows");
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t bytes = n * win * win * sizeof(double);
    const void* da = a;
    const void* db = b;
    if (!is_device_ptr(a)) {
        HIP_TRY(ctx->stage_a.ensure(bytes));
        HIP_TRY(hipMemcpyAsync(ctx->stage_a.p, a, bytes, hipMemcpyHostToDevice, ctx->stream));
        da = ctx->stage_a.p;
    }
    if (b == a) {
        db = da;
    } else if (!is_device_ptr(b)) {
        HIP_TRY(ctx->stage_b.ensure(bytes));
        HIP_TRY(hipMemcpyAsync(ctx->stage_b.p, b, bytes, hipMemcpyHostToDevice, ctx->stream));
        db = ctx->stage_b.p;
    }
    const bool out_dev = is_device_ptr(planes);
    double* dplanes = planes;
    if (!out_dev) {
        HIP_TRY(ctx->stage_out.ensure(bytes));
        dplanes = (double*)ctx->stage_out.p;
    }
    // the stacked windows form an image win wide and n*win high with a
    // one-column grid of step win
    opiv::CorrParams p{};
    p.a = da;
    p.b = db;
    p.pitch = win;
    p.x_start = 0;
    p.y_start = 0;
    p.step_x = (int32_t)win;
    p.step_y = (int32_t)win;
    p.nx = 1;
    p.nx_mul = opiv::fastdiv_mul(1);
    p.first = 0;
    p.n_windows = (int32_t)n;
    p.img_h = (uint32_t)(n * win);
    p.real_compat = variant == OPIV_REAL_COMPAT;
    p.tie_tol = kTieTol;
    p.planes = dplanes;
    if (variant == OPIV_DIRECT_BF16_SPLIT)
        HIP_TRY(opiv::launch_corr_direct(p, (int)win, OPIV_F64, true, ctx->stream));
    else
        HIP_TRY(opiv::launch_corr_fft(p, (int)win, OPIV_F64, true, ctx->stream));
    if (!out_dev) HIP_TRY(hipMemcpyAsync(planes, dplanes, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return OPIV_OK;
}

// device copy of a host input (or the pointer itself when already on the device)
static int stage_in(opiv_ctx* ctx, DevBuf& buf, const void* src, size_t bytes, const void** dev)
{
    if (is_device_ptr(src)) {
        *dev = src;
        return OPIV_OK;
    }
    HIP_TRY(buf.ensure(bytes));
    HIP_TRY(hipMemcpyAsync(buf.p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    *dev = buf.p;
    return OPIV_OK;
}

// frames, forward results and geometry shared by the two adjoint entry points
// the adjoint kernels' view of a plan over device frames and forward results
static opiv::AdjParams adjoint_params(const opiv_plan* plan, const void* da, const void* db,
                                      const opiv_window_result* dfwd)
{
    opiv::AdjParams p{};
    p.a = da;
    p.b = db;
    p.pitch = plan->desc.pitch;
    p.x_start = plan->x_start;
    p.y_start = plan->y_start;
    p.step_x = (int32_t)plan->off_x;
    p.step_y = (int32_t)plan->off_y;
    p.nx = (int32_t)plan->nx;
    p.ny = (int32_t)plan->ny;
    p.n_windows = (int32_t)((size_t)plan->nx * plan->ny);
    p.img_h = (int32_t)plan->desc.img_h;
    p.fwd = dfwd;
    return p;
}

// the adjoint kernels take windows up to 64 x 64
static int adjoint_window_check(const opiv_plan* plan)
{
    if (plan->desc.win > 64)
        return fail(OPIV_E_UNSUPPORTED, "the adjoint takes windows up to 64x64, not %ux%u", plan->desc.win,
                    plan->desc.win);
    return OPIV_OK;
}

static int adjoint_setup(opiv_plan* plan, const void* a, const void* b, const opiv_window_result* fwd,
                         opiv::AdjParams& p)
{
    opiv_ctx* ctx = plan->ctx;
    if (plan->desc.variant == OPIV_REAL_COMPAT)
        return fail(OPIV_E_UNSUPPORTED, "the adjoint supports the circular correlation only");
    if (int rc = adjoint_window_check(plan)) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t count = (size_t)plan->nx * plan->ny;
    const size_t frame_bytes = plan->desc.pitch * plan->desc.img_h * dtype_size(plan->desc.dtype);
    const void *da, *db, *dfwd;
    int rc = stage_in(ctx, ctx->stage_a, a, frame_bytes, &da);
    if (rc == OPIV_OK) rc = stage_in(ctx, ctx->stage_b, b, frame_bytes, &db);
    if (rc == OPIV_OK) rc = check_frames(plan, da, db, "adjoint");
    if (rc != OPIV_OK) return rc;
    if (fwd) {
        rc = stage_in(ctx, ctx->stage_fwd, fwd, count * sizeof(opiv_window_result), &dfwd);
        if (rc != OPIV_OK) return rc;
    } else {
        HIP_TRY(ctx->stage_fwd.ensure(count * sizeof(opiv_window_result)));
        opiv::CorrParams cp = plan_params(plan, da, db, (opiv_window_result*)ctx->stage_fwd.p);
        HIP_TRY(launch_plan(plan, cp, ctx->stream));
        dfwd = ctx->stage_fwd.p;
    }
    p = adjoint_params(plan, da, db, (const opiv_window_result*)dfwd);
    return OPIV_OK;
}

// the VJP's second half: tap weights (into coef, or the context's buffer
// when coef is null) and gradients into device buffers, on the context stream
static int launch_vjp(opiv_plan* plan, opiv::AdjParams& p, const double* d_cot, float* d_ga, float* d_gb,
                      opiv::AdjCoef* coef = nullptr)
{
    opiv_ctx* ctx = plan->ctx;
    if (!coef) {
        HIP_TRY(ctx->adj_coef.ensure((size_t)p.n_windows * sizeof(opiv::AdjCoef)));
        coef = (opiv::AdjCoef*)ctx->adj_coef.p;
    }
    p.coef = coef;
    p.cot = d_cot;
    p.grad_a = d_ga;
    p.grad_b = d_gb;
    p.gpitch = plan->desc.pitch;
    HIP_TRY(opiv::launch_adjoint(p, (int)plan->desc.win, plan->desc.dtype, false, ctx->stream));
    return OPIV_OK;
}

int opiv_vjp_pair_async(opiv_plan* plan, const void* d_a, const void* d_b, const opiv_window_result* d_fwd,
                        const double* d_cot, float* d_grad_a, float* d_grad_b)
{
    if (!plan || !d_a || !d_b || !d_fwd || !d_cot || !d_grad_a || !d_grad_b)
        return fail(OPIV_E_INVALID, "opiv_vjp_pair_async: null argument");
    if (plan->desc.variant == OPIV_REAL_COMPAT)
        return fail(OPIV_E_UNSUPPORTED, "the adjoint supports the circular correlation only");
    if (int rc = adjoint_window_check(plan)) return rc;
    if (int rc = check_frames(plan, d_a, d_b, "opiv_vjp_pair_async")) return rc;
    for (const void* q : { d_a, d_b, (const void*)d_fwd, (const void*)d_cot, (const void*)d_grad_a,
                           (const void*)d_grad_b })
        if (!is_device_ptr(q))
            return fail(OPIV_E_INVALID, "opiv_vjp_pair_async: every buffer must be device memory");
    HIP_TRY(hipSetDevice(plan->ctx->device));
    opiv::AdjParams p = adjoint_params(plan, d_a, d_b, d_fwd);
    return launch_vjp(plan, p, d_cot, d_grad_a, d_grad_b);
}

int opiv_vjp_pair_async_noalloc(opiv_plan* plan, const void* d_a, const void* d_b, const opiv_window_result* d_fwd,
                                const double* d_cot, float* d_grad_a, float* d_grad_b)
{
    if (!plan || !d_a || !d_b || !d_fwd || !d_cot || !d_grad_a || !d_grad_b)
        return fail(OPIV_E_INVALID, "opiv_vjp_pair_async_noalloc: null argument");
    const size_t count = (size_t)plan->nx * plan->ny;
    if (count > OPIV_NOALLOC_MAX_WINDOWS)
        return fail(OPIV_E_UNSUPPORTED, "opiv_vjp_pair_async_noalloc: %zu windows exceed the scratch array's %d",
                    count, OPIV_NOALLOC_MAX_WINDOWS);
    if (plan->desc.variant == OPIV_REAL_COMPAT)
        return fail(OPIV_E_UNSUPPORTED, "the adjoint supports the circular correlation only");
    if (int rc = adjoint_window_check(plan)) return rc;
    if (int rc = check_frames(plan, d_a, d_b, "opiv_vjp_pair_async_noalloc")) return rc;
    for (const void* q : { d_a, d_b, (const void*)d_fwd, (const void*)d_cot, (const void*)d_grad_a,
                           (const void*)d_grad_b })
        if (!is_device_ptr(q))
            return fail(OPIV_E_INVALID, "opiv_vjp_pair_async_noalloc: every buffer must be device memory");
    opiv_ctx* ctx = plan->ctx;
    if (!ctx->scratch_coef)
        return fail(OPIV_E_HIP, "opiv_vjp_pair_async_noalloc: the scratch array of device %d is not mapped",
                    ctx->device);
    HIP_TRY(hipSetDevice(ctx->device));
    opiv::AdjParams p = adjoint_params(plan, d_a, d_b, d_fwd);
    return launch_vjp(plan, p, d_cot, d_grad_a, d_grad_b, static_cast<opiv::AdjCoef*>(ctx->scratch_coef));
}

int opiv_vjp_pair(opiv_plan* plan, const void* a, const void* b, const opiv_window_result* fwd,
                  const double* cot, float* grad_a, float* grad_b)
{
    if (!plan || !a || !b || !cot || !grad_a || !grad_b) return fail(OPIV_E_INVALID, "opiv_vjp_pair: null argument");
    opiv_ctx* ctx = plan->ctx;
    opiv::AdjParams p;
    int rc = adjoint_setup(plan, a, b, fwd, p);
    if (rc != OPIV_OK) return rc;
    const size_t count = (size_t)p.n_windows;
    const void* dcot;
    rc = stage_in(ctx, ctx->stage_cot, cot, count * 2 * sizeof(double), &dcot);
    if (rc != OPIV_OK) return rc;
    const size_t gbytes = plan->desc.pitch * plan->desc.img_h * sizeof(float);
    const bool ga_dev = is_device_ptr(grad_a), gb_dev = is_device_ptr(grad_b);
    float* dga = grad_a;
    float* dgb = grad_b;
    if (!ga_dev) {
        HIP_TRY(ctx->stage_ga.ensure(gbytes));
        dga = (float*)ctx->stage_ga.p;
    }
    if (!gb_dev) {
        HIP_TRY(ctx->stage_gb.ensure(gbytes));
        dgb = (float*)ctx->stage_gb.p;
    }
    rc = launch_vjp(plan, p, (const double*)dcot, dga, dgb);
    if (rc != OPIV_OK) return rc;
    if (!ga_dev) HIP_TRY(hipMemcpyAsync(grad_a, dga, gbytes, hipMemcpyDeviceToHost, ctx->stream));
    if (!gb_dev) HIP_TRY(hipMemcpyAsync(grad_b, dgb, gbytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return OPIV_OK;
}

int opiv_variance_pair(opiv_plan* plan, const void* a, const void* b, const opiv_window_result* fwd,
                       double sigma, double* var)
{
    if (!plan || !a || !b || !var) return fail(OPIV_E_INVALID, "opiv_variance_pair: null argument");
    opiv_ctx* ctx = plan->ctx;
    opiv::AdjParams p;
    int rc = adjoint_setup(plan, a, b, fwd, p);
    if (rc != OPIV_OK) return rc;
    const size_t bytes = (size_t)p.n_windows * 2 * sizeof(double);
    const bool dev = is_device_ptr(var);
    double* dvar = var;
    if (!dev) {
        HIP_TRY(ctx->stage_var.ensure(bytes));
        dvar = (double*)ctx->stage_var.p;
    }
    p.sigma = sigma;
    p.var = dvar;
    HIP_TRY(opiv::launch_adjoint(p, (int)plan->desc.win, plan->desc.dtype, true, ctx->stream));
    if (!dev) HIP_TRY(hipMemcpyAsync(var, dvar, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return OPIV_OK;
}

int opiv_correlate_windows(opiv_ctx* ctx, uint32_t win, size_t n, int32_t variant,
                           const double* a, const double* b, double* planes)
{
    return correlate_stacked(ctx, win, n, variant, a, b, planes);
}

int opiv_auto_correlate_windows(opiv_ctx* ctx, uint32_t win, size_t n, const double* a, double* planes)
{
    // FFT(a) conj(FFT(a)) = |FFT(a)|^2: the cross-correlation of a with itself
    return correlate_stacked(ctx, win, n, OPIV_CIRCULAR_FFT_FP32, a, a, planes);
}

int opiv_average_subtract(opiv_ctx* ctx, const void* const* frames, size_t n, uint32_t img_w, uint32_t img_h,
                          size_t pitch, int32_t dtype, uint16_t* const* out, double* mean)
{
    if (!ctx || !frames || !out) return fail(OPIV_E_INVALID, "average_subtract: null argument");
    // examples/average_subtract/main.cpp:85-89
    if (n < 2) return fail(OPIV_E_INVALID, "require two or more input images");
    if (n > 0x7fffffff) return fail(OPIV_E_UNSUPPORTED, "too many frames");
    if (img_w == 0 || img_h == 0) return fail(OPIV_E_INVALID, "average_subtract: empty frame");
    const size_t esz = dtype_size(dtype);
    if (esz == 0) return fail(OPIV_E_INVALID, "unknown dtype %d", dtype);
    if (pitch == 0) pitch = img_w;
    if (pitch < img_w) return fail(OPIV_E_INVALID, "pitch %zu smaller than width %u", pitch, img_w);
    for (size_t i = 0; i < n; ++i)
        if (!frames[i] || !out[i]) return fail(OPIV_E_INVALID, "average_subtract: null frame or output %zu", i);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t fbytes = pitch * img_h * esz, obytes = (size_t)img_w * img_h * sizeof(uint16_t);
    const size_t mbytes = (size_t)img_w * img_h * sizeof(double);

    // device frame and output pointers; host ones are staged
    std::vector<const void*> dframes(n);
    std::vector<uint16_t*> douts(n);
    std::vector<bool> host_out(n);
    size_t nhost_in = 0, nhost_out = 0;
    for (size_t i = 0; i < n; ++i) {
        nhost_in += is_device_ptr(frames[i]) ? 0 : 1;
        host_out[i] = !is_device_ptr(out[i]);
        nhost_out += host_out[i] ? 1 : 0;
    }
    if (nhost_in) HIP_TRY(ctx->bg_frames.ensure(nhost_in * fbytes));
    if (nhost_out) HIP_TRY(ctx->bg_out.ensure(nhost_out * obytes));
    for (size_t i = 0, hi = 0, ho = 0; i < n; ++i) {
        if (is_device_ptr(frames[i])) {
            dframes[i] = frames[i];
        } else {
            void* dst = static_cast<char*>(ctx->bg_frames.p) + (hi++) * fbytes;
            HIP_TRY(hipMemcpyAsync(dst, frames[i], fbytes, hipMemcpyHostToDevice, ctx->stream));
            dframes[i] = dst;
        }
        douts[i] = host_out[i] ? reinterpret_cast<uint16_t*>(static_cast<char*>(ctx->bg_out.p) + (ho++) * obytes)
                               : out[i];
    }
    const bool mean_dev = mean && is_device_ptr(mean);
    double* dmean = mean_dev ? mean : nullptr;
    if (!dmean) {
        HIP_TRY(ctx->bg_mean.ensure(mbytes));
        dmean = static_cast<double*>(ctx->bg_mean.p);
    }
    HIP_TRY(ctx->bg_table.ensure(2 * n * sizeof(void*)));
    HIP_TRY(ctx->bg_range.ensure(2 * n * sizeof(unsigned long long)));
    std::vector<const void*> table(2 * n);
    for (size_t i = 0; i < n; ++i) {
        table[i] = dframes[i];
        table[n + i] = douts[i];
    }
    HIP_TRY(hipMemcpyAsync(ctx->bg_table.p, table.data(), 2 * n * sizeof(void*), hipMemcpyHostToDevice, ctx->stream));

    opiv::BgParams p{};
    p.frames = static_cast<const void* const*>(ctx->bg_table.p);
    p.out = reinterpret_cast<uint16_t* const*>(static_cast<char*>(ctx->bg_table.p) + n * sizeof(void*));
    p.n = (int)n;
    p.w = (int)img_w;
    p.h = (int)img_h;
    p.pitch = pitch;
    p.mean = dmean;
    p.range = static_cast<unsigned long long*>(ctx->bg_range.p);
    // the run-of-8 kernels: every row start of every frame, output and mean
    // on a 16-byte boundary (8-byte for 8-bit frames, whose runs are 8 bytes)
    const size_t falign = esz == 1 ? 8 : 16;
    bool vec = img_w % 8 == 0 && pitch % 8 == 0 && ((uintptr_t)dmean & 15) == 0;
    for (size_t i = 0; i < n && vec; ++i)
        vec = ((uintptr_t)dframes[i] % falign) == 0 && ((uintptr_t)douts[i] & 15) == 0;
    p.vec = vec;
    HIP_TRY(opiv::launch_average_subtract(p, dtype, ctx->stream));
    for (size_t i = 0; i < n; ++i)
        if (host_out[i]) HIP_TRY(hipMemcpyAsync(out[i], douts[i], obytes, hipMemcpyDeviceToHost, ctx->stream));
    if (mean && !mean_dev) HIP_TRY(hipMemcpyAsync(mean, dmean, mbytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return OPIV_OK;
}

int opiv_fft_transform_windows(opiv_ctx* ctx, uint32_t win, size_t n, int32_t direction,
                               const double* in, double* out)
{
    if (!ctx || !in || !out) return fail(OPIV_E_INVALID, "transform: null argument");
    if (int rc = check_window(win)) return rc;
    if (direction != OPIV_FORWARD && direction != OPIV_REVERSE)
        return fail(OPIV_E_INVALID, "unknown direction %d", direction);
    if (n == 0) return OPIV_OK;
    if (n > 0x7fffffffull / win) return fail(OPIV_E_UNSUPPORTED, "too many windows");
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t bytes = n * win * win * 2 * sizeof(double);
    const double* din = in;
    if (!is_device_ptr(in)) {
        HIP_TRY(ctx->stage_a.ensure(bytes));
        HIP_TRY(hipMemcpyAsync(ctx->stage_a.p, in, bytes, hipMemcpyHostToDevice, ctx->stream));
        din = (const double*)ctx->stage_a.p;
    }
    const bool out_dev = is_device_ptr(out);
    double* dout = out;
    if (!out_dev) {
        HIP_TRY(ctx->stage_out.ensure(bytes));
        dout = (double*)ctx->stage_out.p;
    }
    HIP_TRY(opiv::launch_fft_transform(din, dout, (int)win, n, direction == OPIV_FORWARD, ctx->stream));
    if (!out_dev) HIP_TRY(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return OPIV_OK;
}

} // extern "C"
