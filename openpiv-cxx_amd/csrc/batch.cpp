// batch.cpp -- many frame pairs and several window passes per launch.
//
// Each pass (one window size over every pair) is one launch of the fused
// kernel reading its pairs from a device table, or every pass goes into one
// persistent launch of the multi-pass kernel; either way the batch's work is
// captured once into a HIP graph and replayed by opiv_batch_run.  Launch
// shapes (threads per block, persistent block count per pass) and the launch
// form come from a tuning cache or are timed on the batch's own frames.  The shape never
// changes a window's arithmetic: which windows share a packed inverse FFT is
// fixed by the window numbering, not by the grid, so results are
// bit-identical for every shape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

#include "corr_fft.hpp"
#include "ctx.hpp"
#include "openpiv_amd.h"
#include "plan_check.hpp"

using namespace opiv::detail;

namespace {

struct Shape {
    int nt = 256;       // threads per block
    int blocks = 0;     // 0 = one block per task group; else persistent block count
};

std::string cache_path()
{
    if (const char* p = std::getenv("OPIV_TUNE_CACHE")) return p;
    const char* home = std::getenv("HOME");
    return std::string(home ? home : "/tmp") + "/.cache/openpiv_amd/tune.txt";
}

// cache lines: "<arch> <win> <dtype> <log2 windows> <threads> <blocks per CU x 4>"
std::string cache_key(const std::string& arch, int win, int dtype, size_t nwin)
{
    std::ostringstream k;
    k << arch << " " << win << " " << dtype << " " << (int)std::lround(std::log2((double)std::max<size_t>(1, nwin)));
    return k.str();
}

bool cache_lookup(const std::string& key, int cu, Shape& s)
{
    std::ifstream is(cache_path());
    std::string line;
    while (std::getline(is, line)) {
        if (line.rfind(key + " ", 0) != 0) continue;
        std::istringstream r(line.substr(key.size() + 1));
        int nt = 0, bpc4 = 0;
        // blocks per CU x 4 up to 256 (64 blocks per CU): a bound that keeps
        // the grid arithmetic in range whatever the file holds
        if (r >> nt >> bpc4 && (nt == 64 || nt == 256) && bpc4 >= 0 && bpc4 <= 256) {
            s.nt = nt;
            s.blocks = bpc4 * cu / 4;
            return true;
        }
    }
    return false;
}

// the launch form's line: "<arch> form <dtype> <w0>/<w1>/.. <log2 windows> <single> <blocks per CU x 4>"
std::string form_key(const std::string& arch, const opiv_batch_desc& d, size_t nwin)
{
    std::ostringstream k;
    k << arch << " form " << d.dtype << " ";
    for (uint32_t ps = 0; ps < d.npasses; ++ps) k << (ps ? "/" : "") << d.win[ps];
    k << " " << (int)std::lround(std::log2((double)std::max<size_t>(1, nwin)));
    return k.str();
}

bool cache_lookup_raw(const std::string& key, int& v0, int& v1)
{
    std::ifstream is(cache_path());
    std::string line;
    while (std::getline(is, line)) {
        if (line.rfind(key + " ", 0) != 0) continue;
        std::istringstream r(line.substr(key.size() + 1));
        if (r >> v0 >> v1) return true;
    }
    return false;
}

// replace key's line of the cache with "key value": the file is rewritten
// to a per-process temporary and renamed over the old one, so a reader never
// sees a partial file
void cache_store_line(const std::string& key, const std::string& value)
{
    const std::string path = cache_path();
    std::vector<std::string> keep;
    {
        std::ifstream is(path);
        std::string line;
        while (std::getline(is, line))
            if (!line.empty() && line.rfind(key + " ", 0) != 0) keep.push_back(line);
    }
    for (size_t pos = path.find('/', 1); pos != std::string::npos; pos = path.find('/', pos + 1))
        mkdir(path.substr(0, pos).c_str(), 0755);
    const std::string tmp = path + ".tmp" + std::to_string((long)getpid());
    std::ofstream os(tmp);
    for (const auto& l : keep) os << l << "\n";
    os << key << " " << value << "\n";
    os.close();
    if (os) std::rename(tmp.c_str(), path.c_str());
    else std::remove(tmp.c_str());
}

void cache_store(const std::string& key, int cu, const Shape& s)
{
    cache_store_line(key, std::to_string(s.nt) + " " + std::to_string(cu ? s.blocks * 4 / cu : 0));
}

} // namespace

struct opiv_batch {
    opiv_ctx* ctx = nullptr;
    opiv_batch_desc desc{};
    size_t npairs = 0;
    std::vector<opiv::BatchPair> table;   // [pass][pair]
    std::vector<size_t> offsets;          // npasses * npairs + 1
    std::vector<int32_t> pass_windows;
    std::vector<Shape> shape;
    bool single = false;                  // every pass in one persistent launch (corr_fft_multi.hip)
    int single_blocks = 0;                // its grid
    DevBuf d_table;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    opiv_window_result* graph_out = nullptr;
};

namespace {

// parameters of pass `pass`, writing its results to pass_out
opiv::CorrParams pass_params(const opiv_batch* b, uint32_t pass, opiv_window_result* pass_out)
{
    opiv::CorrParams p{};
    p.n_windows = b->pass_windows[pass];
    p.limit_search = b->desc.limit_search ? 1 : 0;
    p.real_compat = b->desc.variant == OPIV_REAL_COMPAT ? 1 : 0;
    p.tie_tol = kTieTol;
    p.out = pass_out;
    p.pairs = reinterpret_cast<const opiv::BatchPair*>(b->d_table.p) + (size_t)pass * b->npairs;
    p.npairs = (int32_t)b->npairs;
    return p;
}

int fill_table(opiv_batch* b, const opiv_pair_desc* pairs, size_t npairs)
{
    const auto& d = b->desc;
    b->table.assign((size_t)d.npasses * npairs, opiv::BatchPair{});
    b->offsets.assign((size_t)d.npasses * npairs + 1, 0);
    b->pass_windows.assign(d.npasses, 0);
    size_t total = 0;
    for (uint32_t ps = 0; ps < d.npasses; ++ps) {
        const uint32_t win = d.win[ps];
        int rc = OPIV_OK;
        int64_t first = 0;
        const int64_t align = opiv::batch_align((int)win, (int)d.dtype);
        const size_t pass_base = total;
        for (size_t i = 0; i < npairs; ++i) {
            const auto& pr = pairs[i];
            auto& e = b->table[(size_t)ps * npairs + i];
            rc = check_batch_pair(pr, i, ps, win, d.overlap[ps], (int)d.dtype, &e);
            if (rc) return rc;
            if (ps == 0 && (!is_device_ptr(pr.a) || !is_device_ptr(pr.b)))
                return fail(OPIV_E_INVALID, "pair %zu: frames must be device memory", i);
            const uint32_t nx = (uint32_t)e.nx, ny = (uint32_t)e.ny;
            e.first = (int32_t)first;
            e.out_first = (int32_t)(total - pass_base);
            b->offsets[(size_t)ps * npairs + i] = total;
            first += ((int64_t)nx * ny + align - 1) / align * align;   // task-aligned range
            total += (size_t)nx * ny;
            if (first > 0x7fffffff) return fail(OPIV_E_UNSUPPORTED, "pass %u: too many windows", ps);
        }
        b->pass_windows[ps] = (int32_t)first;
    }
    b->offsets.back() = total;
    return OPIV_OK;
}

int upload_table(opiv_batch* b)
{
    const size_t bytes = b->table.size() * sizeof(opiv::BatchPair);
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));   // no graph in flight reads the old table
    HIP_TRY(b->d_table.ensure(bytes));
    HIP_TRY(hipMemcpyAsync(b->d_table.p, b->table.data(), bytes, hipMemcpyHostToDevice, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));   // the host table may change once this returns
    return OPIV_OK;
}

// the two timing events of a tuning run, destroyed on every return path
struct TimingEvents {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t create()
    {
        const hipError_t e = hipEventCreate(&e0);
        return e != hipSuccess ? e : hipEventCreate(&e1);
    }
    ~TimingEvents()
    {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

// time each candidate shape on the batch's own frames and keep the fastest
int tune(opiv_batch* b, uint32_t pass, opiv_window_result* scratch)
{
    const int cu = b->ctx->cu_count > 0 ? b->ctx->cu_count : 256;
    if (b->desc.win[pass] >= 128) {         // one 256-thread workgroup per window: nothing to choose
        b->shape[pass] = Shape{256, 0};
        return OPIV_OK;
    }
    const std::vector<Shape> cands = {{256, 0}, {64, 0}, {256, 2 * cu}, {256, 4 * cu}, {64, 8 * cu}, {64, 16 * cu}};
    const opiv::CorrParams p = pass_params(b, pass, scratch);
    TimingEvents ev;
    HIP_TRY(ev.create());
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    float best = 1e30f;
    Shape pick = cands[0];
    for (const auto& c : cands) {
        HIP_TRY(opiv::launch_corr_fft_shape(p, (int)b->desc.win[pass], b->desc.dtype, false, c.nt, c.blocks,
                                            b->ctx->stream));
        HIP_TRY(hipEventRecord(e0, b->ctx->stream));
        for (int r = 0; r < 3; ++r)
            HIP_TRY(opiv::launch_corr_fft_shape(p, (int)b->desc.win[pass], b->desc.dtype, false, c.nt, c.blocks,
                                                b->ctx->stream));
        HIP_TRY(hipEventRecord(e1, b->ctx->stream));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            pick = c;
        }
    }
    b->shape[pass] = pick;
    return OPIV_OK;
}

// whether the batch has a single-launch form: two passes or more, every
// window 8 to 64
bool single_possible(const opiv_batch* b)
{
    if (b->desc.npasses < 2) return false;
    for (uint32_t ps = 0; ps < b->desc.npasses; ++ps)
        if (b->desc.win[ps] > 64) return false;
    return true;
}

// the single-launch kernel's parameters: pass p writes from out + its offset
opiv::MultiPassParams multi_params(const opiv_batch* b, opiv_window_result* out)
{
    opiv::MultiPassParams mp{};
    mp.npasses = (int32_t)b->desc.npasses;
    for (uint32_t ps = 0; ps < b->desc.npasses; ++ps) {
        mp.pass[ps] = pass_params(b, ps, out + b->offsets[(size_t)ps * b->npairs]);
        mp.win[ps] = (int32_t)b->desc.win[ps];
    }
    return mp;
}

// every pass as the batch launches it: one launch per pass with its shape,
// or the single launch
hipError_t enqueue_passes(const opiv_batch* b, opiv_window_result* out, bool single, int single_blocks,
                          hipStream_t stream)
{
    if (single) return opiv::launch_corr_fft_multipass(multi_params(b, out), b->desc.dtype, single_blocks, stream);
    for (uint32_t ps = 0; ps < b->desc.npasses; ++ps) {
        const hipError_t e = opiv::launch_corr_fft_shape(pass_params(b, ps, out + b->offsets[(size_t)ps * b->npairs]),
                                                         (int)b->desc.win[ps], b->desc.dtype, false,
                                                         b->shape[ps].nt, b->shape[ps].blocks, stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// time the per-pass launches (with their chosen shapes) against the single
// launch at 2 and 4 blocks per CU, back to back on the batch's own frames,
// and keep the fastest form
int tune_form(opiv_batch* b, opiv_window_result* scratch)
{
    const int cu = b->ctx->cu_count > 0 ? b->ctx->cu_count : 256;
    struct Form {
        bool single;
        int blocks;
    };
    const Form forms[] = {{false, 0}, {true, 2 * cu}, {true, 4 * cu}};
    TimingEvents ev;
    HIP_TRY(ev.create());
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    float best = 1e30f;
    Form pick = forms[0];
    for (const auto& f : forms) {
        HIP_TRY(enqueue_passes(b, scratch, f.single, f.blocks, b->ctx->stream));   // warm-up
        HIP_TRY(hipEventRecord(e0, b->ctx->stream));
        for (int r = 0; r < 3; ++r) HIP_TRY(enqueue_passes(b, scratch, f.single, f.blocks, b->ctx->stream));
        HIP_TRY(hipEventRecord(e1, b->ctx->stream));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            pick = f;
        }
    }
    b->single = pick.single;
    b->single_blocks = pick.blocks;
    return OPIV_OK;
}

void drop_graph(opiv_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_out = nullptr;
}

} // namespace

extern "C" {

int opiv_batch_create(opiv_ctx* ctx, const opiv_batch_desc* desc, const opiv_pair_desc* pairs, size_t npairs,
                      opiv_batch** out)
{
    if (!ctx || !desc || !pairs || !out || npairs == 0) return fail(OPIV_E_INVALID, "opiv_batch_create: bad argument");
    *out = nullptr;
    if (desc->npasses < 1 || desc->npasses > 4) return fail(OPIV_E_INVALID, "npasses must be 1..4");
    if (dtype_size(desc->dtype) == 0) return fail(OPIV_E_INVALID, "unknown dtype %d", desc->dtype);
    if (desc->variant != OPIV_CIRCULAR_FFT_FP32 && desc->variant != OPIV_REAL_COMPAT)
        return fail(OPIV_E_UNSUPPORTED, "unknown variant %d", desc->variant);
    for (uint32_t ps = 0; ps < desc->npasses; ++ps)
        if (int rc = check_window(desc->win[ps])) return rc;
    if (npairs > 0x7fffffffull) return fail(OPIV_E_UNSUPPORTED, "too many pairs");
    HIP_TRY(hipSetDevice(ctx->device));
    opiv_batch* b = new opiv_batch;
    b->ctx = ctx;
    b->desc = *desc;
    b->npairs = npairs;
    int rc = fill_table(b, pairs, npairs);
    if (rc == OPIV_OK) rc = upload_table(b);
    if (rc != OPIV_OK) {
        opiv_batch_destroy(b);
        return rc;
    }
    b->shape.assign(desc->npasses, Shape{});
    const char* tune_env = std::getenv("OPIV_TUNE");
    const bool may_time = !(tune_env && tune_env[0] == '0');
    DevBuf scratch;
    for (uint32_t ps = 0; ps < desc->npasses; ++ps) {
        const size_t pass_results = b->offsets[(size_t)(ps + 1) * npairs] - b->offsets[(size_t)ps * npairs];
        const std::string key = cache_key(ctx->arch, (int)desc->win[ps], desc->dtype, pass_results);
        if (cache_lookup(key, ctx->cu_count, b->shape[ps]) || !may_time) continue;
        hipError_t e = scratch.ensure((size_t)b->pass_windows[ps] * sizeof(opiv_window_result));
        if (e != hipSuccess) rc = hip_fail(e, "tuning scratch");
        if (rc == OPIV_OK) rc = tune(b, ps, reinterpret_cast<opiv_window_result*>(scratch.p));
        if (rc != OPIV_OK) break;
        cache_store(key, ctx->cu_count, b->shape[ps]);
    }
    // the launch form: OPIV_BATCH_FORM=single|perpass forces it; otherwise
    // the cache, or timing both forms when tuning is allowed
    if (rc == OPIV_OK && single_possible(b)) {
        const char* form_env = std::getenv("OPIV_BATCH_FORM");
        const std::string fk = form_key(ctx->arch, *desc, b->offsets.back());
        const int cu = ctx->cu_count > 0 ? ctx->cu_count : 256;
        int v0 = 0, v1 = 0;
        if (form_env && std::string(form_env) == "single") {
            b->single = true;
            b->single_blocks = 2 * cu;
        } else if (form_env && std::string(form_env) == "perpass") {
            b->single = false;
        } else if (cache_lookup_raw(fk, v0, v1) && (v0 == 0 || (v0 == 1 && v1 > 0 && v1 <= 256))) {
            b->single = v0 == 1;
            b->single_blocks = v1 * cu / 4;
        } else if (may_time) {
            hipError_t e = scratch.ensure(b->offsets.back() * sizeof(opiv_window_result));
            if (e != hipSuccess) rc = hip_fail(e, "tuning scratch");
            if (rc == OPIV_OK) rc = tune_form(b, reinterpret_cast<opiv_window_result*>(scratch.p));
            if (rc == OPIV_OK)
                cache_store_line(fk, std::to_string(b->single ? 1 : 0) + " " +
                                         std::to_string(b->single ? b->single_blocks * 4 / cu : 0));
        }
    }
    (void)hipStreamSynchronize(ctx->stream);
    scratch.release();
    if (rc != OPIV_OK) {
        opiv_batch_destroy(b);
        return rc;
    }
    *out = b;
    return OPIV_OK;
}

int opiv_batch_destroy(opiv_batch* batch)
{
    if (!batch) return OPIV_OK;
    (void)hipSetDevice(batch->ctx->device);
    (void)hipStreamSynchronize(batch->ctx->stream);
    drop_graph(batch);
    batch->d_table.release();
    delete batch;
    return OPIV_OK;
}

int opiv_batch_window_count(const opiv_batch* batch, size_t* count)
{
    if (!batch || !count) return fail(OPIV_E_INVALID, "opiv_batch_window_count: null argument");
    *count = batch->offsets.back();
    return OPIV_OK;
}

int opiv_batch_offsets(const opiv_batch* batch, size_t* offsets)
{
    if (!batch || !offsets) return fail(OPIV_E_INVALID, "opiv_batch_offsets: null argument");
    std::copy(batch->offsets.begin(), batch->offsets.end(), offsets);
    return OPIV_OK;
}

int opiv_batch_set_pairs(opiv_batch* batch, const opiv_pair_desc* pairs, size_t npairs)
{
    if (!batch || !pairs) return fail(OPIV_E_INVALID, "opiv_batch_set_pairs: null argument");
    if (npairs != batch->npairs) return fail(OPIV_E_SIZE, "pair count %zu differs from the batch's %zu", npairs, batch->npairs);
    std::vector<opiv::BatchPair> old = batch->table;
    std::vector<size_t> old_offsets = batch->offsets;
    int rc = fill_table(batch, pairs, npairs);
    if (rc == OPIV_OK && batch->offsets != old_offsets) rc = fail(OPIV_E_SIZE, "frame sizes differ from the batch's");
    if (rc != OPIV_OK) {
        batch->table = old;
        batch->offsets = old_offsets;
        return rc;
    }
    return upload_table(batch);
}

int opiv_batch_shape(const opiv_batch* batch, uint32_t pass, int32_t* threads, int32_t* blocks)
{
    if (!batch || !threads || !blocks || pass >= batch->desc.npasses)
        return fail(OPIV_E_INVALID, "opiv_batch_shape: bad argument");
    *threads = batch->shape[pass].nt;
    *blocks = batch->shape[pass].blocks;
    return OPIV_OK;
}

int opiv_batch_form(const opiv_batch* batch, int32_t* single, int32_t* blocks)
{
    if (!batch || !single || !blocks) return fail(OPIV_E_INVALID, "opiv_batch_form: null argument");
    *single = batch->single ? 1 : 0;
    *blocks = batch->single ? batch->single_blocks : 0;
    return OPIV_OK;
}

int opiv_batch_run(opiv_batch* batch, opiv_window_result* d_out, size_t n)
{
    if (!batch || !d_out) return fail(OPIV_E_INVALID, "opiv_batch_run: null argument");
    if (n != batch->offsets.back())
        return fail(OPIV_E_SIZE, "result count %zu differs from the batch's %zu windows", n, batch->offsets.back());
    opiv_ctx* ctx = batch->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    if (batch->graph_out != d_out && !is_device_ptr(d_out))
        return fail(OPIV_E_INVALID, "opiv_batch_run: results must be device memory");
    if (!batch->exec || batch->graph_out != d_out) {
        drop_graph(batch);
        HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t le = enqueue_passes(batch, d_out, batch->single, batch->single_blocks, ctx->stream);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
        if (le != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return hip_fail(le, "batch kernel launch during capture");
        }
        HIP_TRY(ce);
        batch->graph = g;
        HIP_TRY(hipGraphInstantiate(&batch->exec, batch->graph, nullptr, nullptr, 0));
        batch->graph_out = d_out;
    }
    HIP_TRY(hipGraphLaunch(batch->exec, ctx->stream));
    return OPIV_OK;
}

} // extern "C"
