// Batched Lloyd k-means over many independent instances resident on the device (MI355X
// extension, no counterpart in the reference; each instance follows the single-instance
// Lloyd iteration of kmeans.cpp:996-1124 with supplied initial centres).
//
// The host builds the work lists (assignment tiles, largest instances first; sample chunks
// for the centroid sums), picks the tile shape and chunk size with a small autotuner that
// times the candidates on the batch itself, and captures the whole run (preparation,
// max_iter iterations, final relabel, inertia, outputs) in one HIP graph. A later call with
// the same batch layout and pointers replays the graph.
#include "../core/gpu.hpp"
#include "../kernels/launch.hpp"
#include "aoclda_amd.h"
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <vector>

namespace {

using namespace dagpu;

struct BatchKey {
    std::vector<int> n, d, k;
    const void *X, *C, *labels;
    int max_iter, prec;
    double tol;
    bool operator<(const BatchKey &o) const {
        return std::tie(n, d, k, X, C, labels, max_iter, prec, tol) <
               std::tie(o.n, o.d, o.k, o.X, o.C, o.labels, o.max_iter, o.prec, o.tol);
    }
};

// Device state of one batch layout.
struct BatchRun {
    GpuContext gpu;
    DeviceBuffer dn, dd, dk, dkp, dxoff, dxboff, dcoff, dcboff, dloff;
    DeviceBuffer xb, c0, c1, l0, l1, cb, chalf, cmax2, tiles, chunks, poff, cbeg, cend,
        list_count, changed, list, escale, part, pcnt, state, fpar, n_iter, cin, amax, inertia;
    KBatchDesc desc{};
    KBatchBuffers w{};
    int n_inst = 0;
    int64_t ctotal = 0;
    int pt = 4, grid = 0;
    size_t lds = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~BatchRun() {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
    }
};

std::mutex g_mu;
std::map<BatchKey, std::unique_ptr<BatchRun>> g_runs;
ErrorStack g_err;
int g_last_tile = 0; // samples per assignment tile of the last batch run

template <typename V> hipError_t upload(DeviceBuffer &b, const std::vector<V> &v, hipStream_t s) {
    hipError_t e = b.reserve(std::max<size_t>(1, v.size()) * sizeof(V));
    if (e != hipSuccess || v.empty())
        return e;
    e = hipMemcpyAsync(b.ptr(), v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice, s);
    return e;
}

// Builds the work lists and buffers of a batch and tunes the tile shape.
// Padded width of an instance's operand copy: the bf16 MFMA takes 32 features per step (so
// 32 or 64, for bf16 and the hi + lo split); the fp32 one 4 per lane group, padded to 8, 16,
// 32 or 64 for vector loads.
int op_width(int prec, int d) {
    if (prec == 1 || prec == 2)
        return d <= 32 ? 32 : 64;
    int kp = 8;
    while (kp < d)
        kp *= 2;
    return kp;
}

da_status build(BatchRun &r, da_int n_inst, const da_int *n, const da_int *d, const da_int *k,
                const float *X, int chunk_hint, int prec) {
    hipStream_t st = r.gpu.stream();
    r.n_inst = n_inst;
    std::vector<int> hn(n, n + n_inst), hd(d, d + n_inst), hk(k, k + n_inst), hkp(n_inst);
    std::vector<int64_t> xoff(n_inst), xboff(n_inst), coff(n_inst), cboff(n_inst), loff(n_inst);
    int64_t xo = 0, xbo = 0, co = 0, cbo = 0, lo = 0;
    const int parts = prec == 2 ? 2 : 1; // operand rows [hi kp][lo kp] for the split form
    for (int i = 0; i < n_inst; i++) {
        hkp[i] = op_width(prec, hd[i]);
        xoff[i] = xo;
        xboff[i] = xbo;
        coff[i] = co;
        cboff[i] = cbo;
        loff[i] = lo;
        xo += (int64_t)hn[i] * hd[i];
        xbo += (int64_t)hn[i] * hkp[i] * parts;
        co += (int64_t)hk[i] * hd[i];
        cbo += (int64_t)((hk[i] + 15) / 16 * 16) * hkp[i] * parts;
        lo += hn[i];
    }
    r.ctotal = co;
    // Sample chunks for sums, maxima and inertia.
    const int cs = chunk_hint;
    std::vector<KChunk> chunks;
    std::vector<int> cbeg(n_inst), cend(n_inst);
    std::vector<int64_t> poff;
    int64_t po = 0;
    for (int i = 0; i < n_inst; i++) {
        cbeg[i] = (int)chunks.size();
        for (int s0 = 0; s0 < hn[i]; s0 += cs) {
            chunks.push_back({i, s0, std::min(cs, hn[i] - s0), 0});
            poff.push_back(po);
            po += (int64_t)hk[i] * hd[i] * 2;
        }
        cend[i] = (int)chunks.size();
    }
    if (chunks.size() > (size_t)INT32_MAX / 2)
        return da_status_invalid_input;
    hipError_t e = hipSuccess;
#define BR(x)                                                                                    \
    do {                                                                                         \
        if ((e = (x)) != hipSuccess)                                                             \
            return hip_status(&g_err, e, #x);                                                    \
    } while (0)
    BR(upload(r.dn, hn, st));
    BR(upload(r.dd, hd, st));
    BR(upload(r.dk, hk, st));
    BR(upload(r.dkp, hkp, st));
    BR(upload(r.dxoff, xoff, st));
    BR(upload(r.dxboff, xboff, st));
    BR(upload(r.dcoff, coff, st));
    BR(upload(r.dcboff, cboff, st));
    BR(upload(r.dloff, loff, st));
    BR(upload(r.chunks, chunks, st));
    BR(upload(r.poff, poff, st));
    BR(upload(r.cbeg, cbeg, st));
    BR(upload(r.cend, cend, st));
    const size_t esz = prec == 0 ? 4 : 2;
    BR(r.xb.reserve((size_t)std::max<int64_t>(1, xbo) * esz));
    BR(r.c0.reserve((size_t)std::max<int64_t>(1, co) * sizeof(double)));
    BR(r.c1.reserve((size_t)std::max<int64_t>(1, co) * sizeof(double)));
    BR(r.l0.reserve((size_t)std::max<int64_t>(1, lo) * sizeof(int)));
    BR(r.l1.reserve((size_t)std::max<int64_t>(1, lo) * sizeof(int)));
    BR(r.cb.reserve((size_t)std::max<int64_t>(1, cbo) * esz));
    BR(r.chalf.reserve((size_t)n_inst * 64 * sizeof(float)));
    BR(r.cmax2.reserve((size_t)n_inst * sizeof(double)));
    BR(r.list_count.reserve(sizeof(int)));
    BR(r.changed.reserve((size_t)n_inst * sizeof(int)));
    BR(r.list.reserve((size_t)std::max<int64_t>(1, lo) * sizeof(KTile)));
    BR(r.escale.reserve((size_t)n_inst * 65 * sizeof(int)));
    BR(r.part.reserve((size_t)std::max<int64_t>(1, po) * sizeof(long long)));
    BR(r.pcnt.reserve(chunks.size() * 64 * sizeof(int)));
    BR(r.state.reserve((size_t)n_inst * sizeof(int)));
    BR(r.fpar.reserve((size_t)n_inst * sizeof(int)));
    BR(r.n_iter.reserve((size_t)n_inst * sizeof(int)));
    BR(r.cin.reserve(chunks.size() * sizeof(double)));
    BR(r.amax.reserve((size_t)n_inst * 64 * sizeof(unsigned long long)));
    BR(r.inertia.reserve((size_t)n_inst * sizeof(double)));
    r.desc = KBatchDesc{r.dn.as<int>(),      r.dd.as<int>(),       r.dk.as<int>(),
                        r.dkp.as<int>(),     r.dxoff.as<int64_t>(), r.dxboff.as<int64_t>(),
                        r.dcoff.as<int64_t>(), r.dcboff.as<int64_t>(), r.dloff.as<int64_t>()};
    KBatchBuffers &w = r.w;
    w.X = X;
    w.prec = prec;
    w.Xo = r.xb.ptr();
    w.C0 = r.c0.as<double>();
    w.C1 = r.c1.as<double>();
    w.L0 = r.l0.as<int>();
    w.L1 = r.l1.as<int>();
    w.Cb = r.cb.ptr();
    w.chalf = r.chalf.as<float>();
    w.cmax2 = r.cmax2.as<double>();
    w.chunks = r.chunks.as<KChunk>();
    w.nchunks = (int)chunks.size();
    w.poff = r.poff.as<int64_t>();
    w.cbeg = r.cbeg.as<int>();
    w.cend = r.cend.as<int>();
    w.list_count = r.list_count.as<int>();
    w.changed = r.changed.as<int>();
    w.list = r.list.as<KTile>();
    w.escale = r.escale.as<int>();
    w.part = r.part.as<long long>();
    w.pcnt = r.pcnt.as<int>();
    w.state = r.state.as<int>();
    w.fpar = r.fpar.as<int>();
    w.n_iter = r.n_iter.as<int>();
    w.cin = r.cin.as<double>();
    // Centroid sums: label offsets and fill counters (130 ints) and the chunk's samples
    // sorted by label, a region that later holds 2,048 per-thread int64 sums.
    int cmax = 1;
    for (const KChunk &c : chunks)
        cmax = std::max(cmax, c.count);
    r.lds = (size_t)(130 + std::max(cmax, 4096)) * sizeof(int);
    BR(kb_set_partial_lds(r.lds));
    BR(hipStreamSynchronize(st));
#undef BR
    return da_status_success;
}

// Assignment tiles of 64*pt samples, instances in decreasing order of per-sample cost
// (k * kp) and size, the tiles of an instance adjacent, so that the persistent grid ends on
// small tiles and a run of dealt tiles mostly shares one instance's centres.
std::vector<KTile> make_tiles(int n_inst, const da_int *n, const da_int *d, const da_int *k,
                              int pt, int prec) {
    std::vector<int> order(n_inst);
    std::iota(order.begin(), order.end(), 0);
    auto cost = [&](int i) { return (int64_t)n[i] * k[i] * op_width(prec, d[i]); };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
    std::vector<KTile> tiles;
    const int sb = 64 * pt;
    for (int i : order)
        for (int s0 = 0; s0 < n[i]; s0 += sb)
            tiles.push_back({i, s0});
    return tiles;
}

da_status set_tiles(BatchRun &r, const std::vector<KTile> &tiles) {
    hipError_t e = upload(r.tiles, tiles, r.gpu.stream());
    if (e == hipSuccess)
        e = hipStreamSynchronize(r.gpu.stream());
    if (e != hipSuccess)
        return hip_status(&g_err, e, "tiles");
    r.w.tiles = r.tiles.as<KTile>();
    r.w.ntiles = (int)tiles.size();
    return da_status_success;
}

da_status enqueue_run(BatchRun &r, const float *Cinit, float *Cout, int *labels, int max_iter,
                      double tol) {
    hipStream_t st = r.gpu.stream();
    hipError_t e = kb_launch_prepare(r.w.X, Cinit, r.desc, r.n_inst, r.w.chunks, r.w.nchunks,
                                     r.ctotal, r.amax.as<unsigned long long>(), r.w.escale,
                                     r.w.prec, r.w.Xo, r.w.C0, r.w.state, r.w.fpar, r.w.n_iter,
                                     st);
    for (int it = 0; it < max_iter && e == hipSuccess; it++) {
        const int par = it & 1;
        e = kb_launch_assign(r.w, r.desc, r.n_inst, 0, par, r.pt, r.grid, st);
        if (e == hipSuccess)
            e = kb_launch_update(r.w, r.desc, r.n_inst, par, it, tol, r.lds, st);
    }
    if (e == hipSuccess)
        e = kb_launch_finish(r.w, r.desc, r.n_inst, r.pt, r.grid, Cout, labels,
                             r.inertia.as<double>(), st);
    if (e != hipSuccess)
        return hip_status(&g_err, e, "batched k-means launch");
    return da_status_success;
}

// Autotuner: times one assignment pass for each tile shape on the batch (the pass is
// idempotent: it writes the labels from the current centres) and keeps the fastest.
da_status tune(BatchRun &r, const da_int *n, const da_int *d, const da_int *k, const float *Cinit) {
    hipStream_t st = r.gpu.stream();
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    hipError_t e = kb_launch_prepare(r.w.X, Cinit, r.desc, r.n_inst, r.w.chunks, r.w.nchunks,
                                     r.ctotal, r.amax.as<unsigned long long>(), r.w.escale,
                                     r.w.prec, r.w.Xo, r.w.C0, r.w.state, r.w.fpar, r.w.n_iter,
                                     st);
    if (e != hipSuccess)
        return hip_status(&g_err, e, "batched k-means preparation");
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    float best = 0;
    int best_pt = 4;
    for (int pt : {2, 4}) {
        std::vector<KTile> tiles = make_tiles(r.n_inst, n, d, k, pt, r.w.prec);
        da_status s = set_tiles(r, tiles);
        if (s != da_status_success)
            return s;
        const int grid = std::max(1, cus * kb_assign_blocks_per_cu(r.w.prec, pt));
        (void)kb_launch_assign(r.w, r.desc, r.n_inst, 0, 0, pt, grid, st); // warm-up
        (void)hipEventRecord(a, st);
        for (int rep = 0; rep < 3; rep++)
            (void)kb_launch_assign(r.w, r.desc, r.n_inst, 0, 0, pt, grid, st);
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        if (best == 0 || ms < best) {
            best = ms;
            best_pt = pt;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    r.pt = best_pt;
    r.grid = std::max(1, cus * kb_assign_blocks_per_cu(r.w.prec, best_pt));
    return set_tiles(r, make_tiles(r.n_inst, n, d, k, best_pt, r.w.prec));
}

} // namespace

extern "C" {

da_status da_kmeans_batch_device_s(da_int n_inst, const da_int *n, const da_int *d,
                                   const da_int *k, const float *X, float *C, da_int *labels,
                                   da_int max_iter, float tol, da_int cross_term,
                                   da_int *n_iter, float *inertia) {
    if (n_inst < 1)
        return da_status_invalid_array_dimension;
    if (!n || !d || !k || !X || !C || !labels || !n_iter || !inertia)
        return da_status_invalid_pointer;
    if (max_iter < 1 || !(tol >= 0.0f) || cross_term < 0 || cross_term > 2)
        return da_status_invalid_input;
    for (da_int i = 0; i < n_inst; i++) {
        if (n[i] < 1 || d[i] < 1 || d[i] > 64 || k[i] < 1 || k[i] > 64 || k[i] > n[i])
            return da_status_invalid_array_dimension;
    }
    std::lock_guard<std::mutex> g(g_mu);
    BatchKey key{std::vector<int>(n, n + n_inst), std::vector<int>(d, d + n_inst),
                 std::vector<int>(k, k + n_inst), X, C, labels, (int)max_iter, (int)cross_term,
                 (double)tol};
    auto it = g_runs.find(key);
    BatchRun *r = nullptr;
    if (it == g_runs.end()) {
        auto run = std::make_unique<BatchRun>();
        da_status s = run->gpu.init(&g_err);
        if (s != da_status_success)
            return s;
        // Chunks of about 16 K samples (the centroid partials of a chunk are k*d*16 bytes).
        s = build(*run, n_inst, n, d, k, X, 16384, (int)cross_term);
        if (s != da_status_success)
            return s;
        s = tune(*run, n, d, k, C);
        if (s != da_status_success)
            return s;
        hipStream_t st = run->gpu.stream();
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess)
            return hip_status(&g_err, e, "hipStreamBeginCapture");
        s = enqueue_run(*run, C, C, labels, max_iter, tol);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(st, &graph);
        if (s != da_status_success) {
            if (graph)
                (void)hipGraphDestroy(graph);
            return s;
        }
        if (e != hipSuccess)
            return hip_status(&g_err, e, "hipStreamEndCapture");
        run->graph = graph;
        e = hipGraphInstantiate(&run->exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess)
            return hip_status(&g_err, e, "hipGraphInstantiate");
        r = run.get();
        g_runs[key] = std::move(run);
    } else {
        r = it->second.get();
    }
    g_last_tile = 64 * r->pt;
    hipStream_t st = r->gpu.stream();
    hipError_t e = hipGraphLaunch(r->exec, st);
    if (e != hipSuccess)
        return hip_status(&g_err, e, "hipGraphLaunch");
    std::vector<int> it_h(n_inst);
    std::vector<double> in_h(n_inst);
    e = hipMemcpyAsync(it_h.data(), r->n_iter.ptr(), n_inst * sizeof(int), hipMemcpyDeviceToHost,
                       st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(in_h.data(), r->inertia.ptr(), n_inst * sizeof(double),
                           hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return hip_status(&g_err, e, "batched k-means results");
    for (da_int i = 0; i < n_inst; i++) {
        n_iter[i] = it_h[i];
        inertia[i] = (float)in_h[i];
    }
    return da_status_success;
}

da_status da_kmeans_batch_tile_shape(da_int *samples_per_tile) {
    if (!samples_per_tile)
        return da_status_invalid_pointer;
    std::lock_guard<std::mutex> g(g_mu);
    *samples_per_tile = g_last_tile;
    return da_status_success;
}

da_status da_kmeans_batch_release(void) {
    std::lock_guard<std::mutex> g(g_mu);
    g_runs.clear();
    return da_status_success;
}

} // extern "C"
