// pfc_batch.hip -- the pfc_batch_* C ABI: many independent systems of uneven
// sizes stepped together.  Systems of one (Nx, Ny, U) are cut into chunks of a
// cache-sized working set; each chunk's passes are launches of one block per
// work item of all its systems (pfc_batch.hpp), chained step after step on one
// of a few streams; blocks of steps are captured as HIP graphs; a small
// autotuner picks the FFT shape of every (length, U) and the chunk budget.
//
// Measured on the batched-relaxation workload (96 systems, 256^2 to 2048^2,
// mixed precision): one launch per class over all its systems and a join after
// every pass ran 1.88 ms per step; four streams 1.65; the same systems stepped
// one by one with a join per step 1.65 and unsynchronised 1.49; 20-step graph
// blocks and 16 MiB chunks (the row transforms of a chunk are still in the
// Infinity Cache when its column pass reads them) bring the batch to 1.55;
// one block per work item instead of persistent blocks walking the items with
// a grid stride, 1.50 against 1.67 on the same box.  Later, on other boxes:
// 1.36 against 1.42 one by one; descriptors read through the constant address
// space (no spills left in the FP32 kernels) 1.25 against 1.42; the third,
// half-block candidate and the stream policy in the tuning 1.27 against 1.47.
// By class, the batch is 2-4x faster than one by one for sides up to 512 and
// within 10 % either way for 1536 and 2048 (three-field 2048^2 systems run
// 0.42 ms batched against 0.39 one by one; running those one after another on
// one stream (policy 1) made the whole batch slower, 1.35 ms).
//
// A batch borrows existing contexts: their buffers, parameters, box tables and
// NaN flags.  The contexts stay usable through the one-system calls (fields
// in and out, evaluate, box); a batch step advances all of them by the same
// steps, with the refresh of every 100th step as pfc_step does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/pfc_relax.h"
#include "pfc_batch.hpp"

extern template struct pfc::BatchLaunch<float>;
extern template struct pfc::BatchLaunch<double>;

namespace pfc {

__global__ void k_batch_tick(int* tdev, int nsteps) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *tdev += nsteps;
}

namespace {

// the k-th candidate class of (length n, fields U), the default first; -1
// past the last
int class_of(int n, int U, int k) {
    int found = 0;
    for (int c = 0; c < kBatchClasses; c++)
        if (kBatchClass[c].n == n && kBatchClass[c].u == U && found++ == k) return c;
    return -1;
}

// autotuner results per (precision, pass, length, U), for the process
std::map<std::tuple<int, int, int, int>, int>& tuned() {
    static std::map<std::tuple<int, int, int, int>, int> m;
    return m;
}

}  // namespace
}  // namespace pfc

using namespace pfc;

// The launches of one pass: per chunk, its class and its items' offset and
// count (one block per item), largest chunks first.
struct BatchPass {
    std::vector<int> cls, off, cnt;
    std::vector<double> work;
    BatchItem* d = nullptr;
};

// Streams the classes of a pass are spread over (the hardware queues a
// process gets by default), so that small classes overlap large ones.
constexpr int kBatchStreams = 4;
// steps per captured block (a divisor of the refresh period 100)
constexpr int kBatchBlock = 20;
static_assert(100 % kBatchBlock == 0, "a refresh must fall on a block's first step");

struct pfc_batch {
    int n = 0, prec = 0, flags = 0;
    std::vector<pfc_ctx*> ctx;
    hipStream_t s = nullptr;
    hipStream_t ss[kBatchStreams] = {};
    hipEvent_t fork = nullptr, join[kBatchStreams] = {};
    void* dsys = nullptr;    // BatchSys<T>[n]
    void* dtune = nullptr;   // BatchSys<T>[n] with Nh, I redirected to scratch (tuning)
    BatchPass rows, cols;
    std::vector<int> gstream;   // stream of each chunk (entry) of a step
    double chunk_bytes = 16.0 * (1 << 20);
    bool chunk_fixed = false;
    // stream policy: 0 spreads every chunk over the streams by work; 1 runs
    // the chunks larger than the budget (systems that alone exceed it) one
    // after another on stream 0 and spreads the rest over the other streams
    int policy = 0;
    bool policy_fixed = false;
    int* tdev = nullptr;
    int* tnan = nullptr;      // NaN flag of the tuning passes
    double* tn0 = nullptr;    // k = 0 states of the tuning passes (8 doubles per system)
    int thost = 0;
    hipGraph_t graph[4] = {};        // 1 step, 1 refresh step, block, block from a refresh
    hipGraphExec_t exec[4] = {};
    std::map<std::pair<int, int>, int> rclass, cclass;   // (length, U) -> class
};

namespace {

// A system's descriptor; the scratch variant (autotuning) reads and writes
// the context's evaluate buffers R (as I) and O (as Nh) and reports NaNs to a
// scratch flag, so timing passes change no state.
template <typename T>
BatchSys<T> sys_of(const pfc_ctx* c, bool scratch, int* scratch_nan, double* scratch_n0) {
    BatchSys<T> b{};
    b.I = (cpx<T>*)(scratch ? c->R : c->I);
    b.J = (cpx<T>*)c->J;
    b.K = (cpx<T>*)c->K;
    b.Nh = (cpx<T>*)(scratch ? c->O : c->Nh);
    b.fstride = c->nk;
    b.nstride = c->nkn;
    b.H = c->Nx / 2;
    b.sI = c->sI;
    b.sJ = c->sJ;
    b.twx = (const cpx<T>*)c->twx;
    b.twy = (const cpx<T>*)c->twy;
    b.ktab = (const T*)c->ktab;
    b.prm = c->dprm;
    b.nan = scratch ? scratch_nan : c->nan_flag;
    b.n0 = scratch ? scratch_n0 : c->n0;
    return b;
}

// Items of one pass for the systems in `which` (all when empty) under the
// class choice `classes`: one launch entry per chunk of systems with the same
// (Nx, Ny, U), larger chunks first (the same order in both passes, so entry k
// of the row and of the column pass are the same systems).  Each entry's
// items are permuted so that the blocks of one XCD (block b runs on XCD
// b % 8) take consecutive items.
template <typename T>
int build_pass(const pfc_batch* b, bool row_pass, const std::map<std::pair<int, int>, int>& classes,
               const std::vector<int>& which, BatchPass* out) {
    std::vector<int> sel = which;
    if (sel.empty())
        for (int i = 0; i < b->n; i++) sel.push_back(i);
    // groups of equal (Nx, Ny, U), cut into chunks whose row-transform
    // buffers (written by the row pass, read back by the column pass) stay
    // within the chunk budget, so they are still cached when read back
    std::map<std::tuple<int, int, int>, std::vector<int>> groups;
    for (int i : sel) {
        const pfc_ctx* c = b->ctx[i];
        groups[std::make_tuple(c->Nx, c->Ny, c->U)].push_back(i);
    }
    std::vector<std::pair<double, std::vector<int>>> order;
    for (auto& g : groups) {
        std::vector<int> chunk;
        double bytes = 0.0;
        for (int i : g.second) {
            const pfc_ctx* c = b->ctx[i];
            const double sb = (double)c->U * c->nk * c->esz;
            if (!chunk.empty() && bytes + sb > b->chunk_bytes) {
                order.push_back({bytes, chunk});
                chunk.clear();
                bytes = 0.0;
            }
            chunk.push_back(i);
            bytes += sb;
        }
        if (!chunk.empty()) order.push_back({bytes, chunk});
    }
    std::stable_sort(order.begin(), order.end(),
                     [](const auto& x, const auto& y) { return x.first > y.first; });
    std::vector<BatchItem> all;
    BatchPass p;
    for (auto& og : order) {
        const pfc_ctx* c0 = b->ctx[og.second[0]];
        const int Nx = c0->Nx, Ny = c0->Ny, U = c0->U;
        const int cls = classes.at({row_pass ? Nx : Ny, U});
        std::vector<BatchItem> v;
        for (int i : og.second) {
            if (row_pass) {
                for (int g = 0; g < Ny / batch_tr<T>(cls); g++) v.push_back({i, g});
            } else {
                for (int cb = 0; cb <= Nx / 2 / batch_tc<T>(cls); cb++) v.push_back({i, cb});
            }
        }
        const int n8 = (int)v.size() / 8 * 8;
        if (n8 >= 8) {
            std::vector<BatchItem> q(v);
            for (int w = 0; w < n8; w++) q[w] = v[(w % 8) * (n8 / 8) + w / 8];
            v.swap(q);
        }
        p.cls.push_back(cls);
        p.off.push_back((int)all.size());
        p.cnt.push_back((int)v.size());
        p.work.push_back(og.first);
        all.insert(all.end(), v.begin(), v.end());
    }
    if (hipMalloc((void**)&p.d, std::max<size_t>(1, all.size()) * sizeof(BatchItem)) != hipSuccess)
        return PFC_ENOMEM;
    if (hipMemcpy(p.d, all.data(), all.size() * sizeof(BatchItem), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p.d);
        return PFC_EDEVICE;
    }
    if (out->d) (void)hipFree(out->d);
    *out = p;
    return PFC_OK;
}

template <typename T>
void launch_entry(const pfc_batch* b, bool row_pass, const BatchPass& p, int k, bool ref,
                  const void* sys, hipStream_t st, int toff = 0) {
    if (row_pass)
        BatchLaunch<T>::row(p.cls[k], ref, p.d + p.off[k], p.cnt[k], (const BatchSys<T>*)sys,
                            b->tdev, toff, st);
    else
        BatchLaunch<T>::col(p.cls[k], ref, p.d + p.off[k], p.cnt[k], (const BatchSys<T>*)sys, st);
}

// Forks the first ns batch streams from the batch stream / joins them back.
void fork_streams(const pfc_batch* b, int ns) {
    (void)hipEventRecord(b->fork, b->s);
    for (int i = 0; i < ns; i++) (void)hipStreamWaitEvent(b->ss[i], b->fork, 0);
}
void join_streams(const pfc_batch* b, int ns) {
    for (int i = 0; i < ns; i++) {
        (void)hipEventRecord(b->join[i], b->ss[i]);
        (void)hipStreamWaitEvent(b->s, b->join[i], 0);
    }
}

// nsteps steps (the first a refresh step if ref): each chunk's row pass then
// column pass, step after step, on the chunk's stream (chunks spread over the
// streams by work), so a chunk never waits for another; the streams join
// once, at the end, before the step counter advances.
template <typename T>
void launch_steps(const pfc_batch* b, bool ref, int nsteps, const void* rsys = nullptr,
                  const void* csys = nullptr) {
    if (!rsys) rsys = b->dsys;
    if (!csys) csys = b->dsys;
    const int ng = (int)b->rows.cls.size();
    const int ns = std::min<int>(kBatchStreams, ng);
    if (ns > 1) fork_streams(b, ns);
    for (int k = 0; k < ng; k++) {
        hipStream_t st = ns > 1 ? b->ss[b->gstream[k]] : b->s;
        for (int i = 0; i < nsteps; i++) {
            launch_entry<T>(b, true, b->rows, k, ref && i == 0, rsys, st, i);
            launch_entry<T>(b, false, b->cols, k, ref && i == 0, csys, st);
        }
    }
    if (ns > 1) join_streams(b, ns);
    hipLaunchKernelGGL(k_batch_tick, dim3(1), dim3(64), 0, b->s, b->tdev, nsteps);
}

// One pass alone (the autotuner's timing): entries round-robin over the
// streams, joined at the end.
template <typename T>
void launch_pass(const pfc_batch* b, bool row_pass, const BatchPass& p, bool ref, const void* sys) {
    const int ns = std::min<int>(kBatchStreams, (int)p.cls.size());
    if (ns > 1) fork_streams(b, ns);
    for (size_t k = 0; k < p.cls.size(); k++)
        launch_entry<T>(b, row_pass, p, (int)k, ref, sys, ns > 1 ? b->ss[k % ns] : b->s);
    if (ns > 1) join_streams(b, ns);
}

// Best of 3 timed launches (after one untimed) of one pass over the systems in
// `which` with class choice `classes`, in ms; 1e30 on error.  The passes run
// on the scratch descriptors, so no context's state changes.
template <typename T>
float time_pass(pfc_batch* b, bool row_pass, const std::map<std::pair<int, int>, int>& classes,
                const std::vector<int>& which) {
    BatchPass p;
    if (build_pass<T>(b, row_pass, classes, which, &p) != PFC_OK) return 1e30f;
    hipEvent_t e0, e1;
    float best = 1e30f;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        for (int rep = 0; rep < 4; rep++) {
            (void)hipEventRecord(e0, b->s);
            launch_pass<T>(b, row_pass, p, false, b->dtune);
            (void)hipEventRecord(e1, b->s);
            float ms = 1e30f;
            if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
                rep > 0)
                best = std::min(best, ms);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    (void)hipFree(p.d);
    if (hipGetLastError() != hipSuccess) return 1e30f;
    return best;
}

// The launch entries of both passes under the current classes and chunk
// budget, and each chunk's stream: largest first onto the least loaded.
template <typename T>
int build_step(pfc_batch* b) {
    int rc = build_pass<T>(b, true, b->rclass, {}, &b->rows);
    if (rc == PFC_OK) rc = build_pass<T>(b, false, b->cclass, {}, &b->cols);
    if (rc != PFC_OK) return rc;
    double load[kBatchStreams] = {};
    b->gstream.clear();
    for (double w : b->rows.work) {
        int i;
        if (b->policy == 1 && w > b->chunk_bytes) i = 0;
        else {
            const int lo = b->policy == 1 ? 1 : 0;
            i = (int)(std::min_element(load + lo, load + kBatchStreams) - load);
        }
        b->gstream.push_back(i);
        load[i] += w;
    }
    return PFC_OK;
}

// Best of 3 timings (after one untimed) of 4 whole steps under the current
// chunk budget, in ms, on the scratch descriptors (no context's state changes).
template <typename T>
float time_steps(pfc_batch* b) {
    if (build_step<T>(b) != PFC_OK) return 1e30f;
    hipEvent_t e0, e1;
    float best = 1e30f;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        for (int rep = 0; rep < 4; rep++) {
            (void)hipEventRecord(e0, b->s);
            launch_steps<T>(b, false, 4, b->dtune, b->dtune);
            (void)hipEventRecord(e1, b->s);
            float ms = 1e30f;
            if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
                rep > 0)
                best = std::min(best, ms);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    if (hipGetLastError() != hipSuccess) return 1e30f;
    return best;
}

// Chooses the class of every (length, U) of one pass: the tuned one when this
// process knows it, else by timing both candidates (tune) or the first.
template <typename T>
void choose_classes(pfc_batch* b, bool row_pass, bool tune) {
    auto& classes = row_pass ? b->rclass : b->cclass;
    std::map<std::pair<int, int>, std::vector<int>> groups;
    for (int i = 0; i < b->n; i++) {
        const pfc_ctx* c = b->ctx[i];
        groups[{row_pass ? c->Nx : c->Ny, c->U}].push_back(i);
    }
    // candidates whose kernel spills are dropped where a candidate without
    // spills exists; the first remaining one is the default
    // (candidates with the same tile as an earlier one, as halving a single
    // column per block gives, are skipped)
    auto candidates = [&](int n, int U) {
        std::vector<int> all, clean;
        for (int k = 0; k < kBatchCandidates; k++) {
            const int c = class_of(n, U, k);
            if (c < 0) continue;
            bool dup = false;
            for (int a : all)
                dup = dup || (kBatchClass[a].p == kBatchClass[c].p &&
                              (row_pass ? batch_tr<T>(a) == batch_tr<T>(c) : batch_tc<T>(a) == batch_tc<T>(c)));
            if (dup) continue;
            all.push_back(c);
            const int sc = row_pass ? BatchLaunch<T>::row_scratch(c) : BatchLaunch<T>::col_scratch(c);
            if (sc == 0) clean.push_back(c);
        }
        return clean.empty() ? all : clean;
    };
    for (auto& g : groups) classes[g.first] = candidates(g.first.first, g.first.second)[0];
    // PFC_BATCH_CANDIDATE=k: the k-th candidate of the class table wherever
    // it exists (spilling or not, no tuning), for tests of every kernel shape
    if (const char* e = getenv("PFC_BATCH_CANDIDATE")) {
        const int k = atoi(e);
        for (auto& g : groups) {
            const int c = class_of(g.first.first, g.first.second, k);
            if (c >= 0) classes[g.first] = c;
        }
        return;
    }
    for (auto& g : groups) {
        const auto key = std::make_tuple(b->prec, row_pass ? 0 : 1, g.first.first, g.first.second);
        auto hit = tuned().find(key);
        if (hit != tuned().end()) {
            classes[g.first] = hit->second;
            continue;
        }
        if (!tune) continue;
        const std::vector<int> cand = candidates(g.first.first, g.first.second);
        if (cand.size() < 2) {
            tuned()[key] = classes[g.first];
            continue;
        }
        float best = 1e30f;
        for (int cls : cand) {
            std::map<std::pair<int, int>, int> trial = classes;
            trial[g.first] = cls;
            const float t = time_pass<T>(b, row_pass, trial, g.second);
            if (t < best) {
                best = t;
                classes[g.first] = trial[g.first];
            }
        }
        tuned()[key] = classes[g.first];
    }
}

template <typename T>
int build(pfc_batch* b) {
    std::vector<BatchSys<T>> hs(b->n), ht(b->n);
    for (int i = 0; i < b->n; i++) {
        hs[i] = sys_of<T>(b->ctx[i], false, nullptr, nullptr);
        ht[i] = sys_of<T>(b->ctx[i], true, b->tnan, b->tn0 + 8 * i);
    }
    if (hipMalloc(&b->dsys, b->n * sizeof(BatchSys<T>)) != hipSuccess ||
        hipMalloc(&b->dtune, b->n * sizeof(BatchSys<T>)) != hipSuccess)
        return PFC_ENOMEM;
    if (hipMemcpy(b->dsys, hs.data(), b->n * sizeof(BatchSys<T>), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->dtune, ht.data(), b->n * sizeof(BatchSys<T>), hipMemcpyHostToDevice) != hipSuccess)
        return PFC_EDEVICE;
    const bool tune = (b->flags & PFC_BATCH_AUTOTUNE) != 0;
    choose_classes<T>(b, true, tune);
    choose_classes<T>(b, false, tune);
    // chunk budget and stream policy: the environment's, else timed (tune),
    // else the defaults
    if (tune && !(b->chunk_fixed && b->policy_fixed)) {
        float best = 1e30f;
        double pick = b->chunk_bytes;
        int pol = b->policy;
        std::vector<double> budgets = {8.0, 16.0, 32.0, 64.0};
        if (b->chunk_fixed) budgets = {b->chunk_bytes / (1 << 20)};
        std::vector<int> policies = {0, 1};
        if (b->policy_fixed) policies = {b->policy};
        for (int po : policies)
            for (double mb : budgets) {
                b->chunk_bytes = mb * (1 << 20);
                b->policy = po;
                const float t = time_steps<T>(b);
                if (t < best) {
                    best = t;
                    pick = b->chunk_bytes;
                    pol = po;
                }
            }
        b->chunk_bytes = pick;
        b->policy = pol;
    }
    if (getenv("PFC_BATCH_VERBOSE")) {
        fprintf(stderr, "pfc_batch: %d systems, chunk budget %.0f MiB, stream policy %d\n", b->n,
                b->chunk_bytes / (1 << 20), b->policy);
        for (int pass = 0; pass < 2; pass++)
            for (auto& e : pass == 0 ? b->rclass : b->cclass)
                fprintf(stderr, "pfc_batch:   %s length %d U=%d: %d points per thread, %d per block\n",
                        pass == 0 ? "rows" : "columns", e.first.first, e.first.second,
                        kBatchClass[e.second].p, pass == 0 ? batch_tr<T>(e.second) : batch_tc<T>(e.second));
    }
    const int rc = build_step<T>(b);
    if (rc != PFC_OK) return rc;
    if (hipGetLastError() != hipSuccess) return PFC_EDEVICE;
    if (b->flags & PFC_BATCH_NO_GRAPH) return PFC_OK;
    // graphs of one step and of kBatchBlock steps, each without and with a
    // refresh on its first step
    for (int g = 0; g < 4; g++) {
        if (hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return PFC_EDEVICE;
        launch_steps<T>(b, (g & 1) != 0, g < 2 ? 1 : kBatchBlock);
        if (hipStreamEndCapture(b->s, &b->graph[g]) != hipSuccess ||
            hipGraphInstantiate(&b->exec[g], b->graph[g], nullptr, nullptr, 0) != hipSuccess)
            return PFC_EDEVICE;
    }
    return PFC_OK;
}

template <typename T>
void step_eager(pfc_batch* b, bool ref, int nsteps) {
    launch_steps<T>(b, ref, nsteps);
}

bool batch_length(int n) { return class_of(n, 1, 0) >= 0; }

}  // namespace

extern "C" {

int pfc_batch_destroy(pfc_batch* b) {
    if (!b) return PFC_EINVAL;
    if (b->s) (void)hipStreamSynchronize(b->s);
    for (int i = 0; i < 4; i++) {
        if (b->exec[i]) (void)hipGraphExecDestroy(b->exec[i]);
        if (b->graph[i]) (void)hipGraphDestroy(b->graph[i]);
    }
    if (b->dsys) (void)hipFree(b->dsys);
    if (b->dtune) (void)hipFree(b->dtune);
    if (b->rows.d) (void)hipFree(b->rows.d);
    if (b->cols.d) (void)hipFree(b->cols.d);
    if (b->tdev) (void)hipFree(b->tdev);
    if (b->tnan) (void)hipFree(b->tnan);
    if (b->tn0) (void)hipFree(b->tn0);
    for (int i = 0; i < kBatchStreams; i++) {
        if (b->ss[i]) (void)hipStreamDestroy(b->ss[i]);
        if (b->join[i]) (void)hipEventDestroy(b->join[i]);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->s) (void)hipStreamDestroy(b->s);
    delete b;
    return PFC_OK;
}

int pfc_batch_create(pfc_batch** out, pfc_ctx* const* ctxs, int n, int flags) {
    if (!out || !ctxs || n < 1) return PFC_EINVAL;
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        const pfc_ctx* c = ctxs[i];
        if (!c || c->prec != ctxs[0]->prec) return PFC_EINVAL;
        if (c->U != 1 && c->U != 3) return PFC_EINVAL;
        if (c->generic) return PFC_ESIZE;
        if (!batch_length(c->Nx) || !batch_length(c->Ny)) return PFC_ESIZE;
        if (c->sI != ilog2(c->Ny) || c->sJ != ilog2(c->Ny)) return PFC_EINVAL;
    }
    pfc_batch* b = new pfc_batch();
    b->n = n;
    b->prec = ctxs[0]->prec;
    b->flags = flags;
    b->ctx.assign(ctxs, ctxs + n);
    if (const char* e = getenv("PFC_BATCH_CHUNK_MB")) {
        b->chunk_bytes = atof(e) * (1 << 20);
        b->chunk_fixed = true;
    }
    if (const char* e = getenv("PFC_BATCH_POLICY")) {
        b->policy = atoi(e) == 1 ? 1 : 0;
        b->policy_fixed = true;
    }
    if (hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipMalloc((void**)&b->tdev, sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&b->tnan, sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&b->tn0, 8 * sizeof(double) * n) != hipSuccess) {
        pfc_batch_destroy(b);
        return PFC_EDEVICE;
    }
    for (int i = 0; i < kBatchStreams; i++)
        if (hipStreamCreateWithFlags(&b->ss[i], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) != hipSuccess) {
            pfc_batch_destroy(b);
            return PFC_EDEVICE;
        }
    for (pfc_ctx* c : b->ctx)
        if (hipStreamSynchronize(c->stream) != hipSuccess) {
            pfc_batch_destroy(b);
            return PFC_EDEVICE;
        }
    const int rc = (b->prec == PFC_FP64) ? build<double>(b) : build<float>(b);
    if (rc != PFC_OK || hipStreamSynchronize(b->s) != hipSuccess) {
        pfc_batch_destroy(b);
        return rc != PFC_OK ? rc : PFC_EDEVICE;
    }
    *out = b;
    return PFC_OK;
}

int pfc_batch_step(pfc_batch* b, int t0, int nsteps) {
    if (!b || nsteps < 0) return PFC_EINVAL;
    for (pfc_ctx* c : b->ctx) {
        if (!c->have_fields || !c->have_box) return PFC_ESTATE;
        if (hipStreamSynchronize(c->stream) != hipSuccess) return PFC_EDEVICE;
    }
    // the batch refreshes at every t % 100 step itself; after it a
    // context's state is no longer the one its last evaluate computed
    if (nsteps > 0)
        for (pfc_ctx* c : b->ctx) c->nh_fresh = 0;
    b->thost = t0;
    if (hipMemcpyAsync(b->tdev, &b->thost, sizeof(int), hipMemcpyHostToDevice, b->s) != hipSuccess)
        return PFC_EDEVICE;
    // blocks of kBatchBlock steps where they start on a multiple of it (it
    // divides 100, so a refresh can only be a block's first step), single
    // steps elsewhere
    for (int t = t0; t < t0 + nsteps;) {
        const bool ref = (t % 100 == 0);
        const int len = (t % kBatchBlock == 0 && t + kBatchBlock <= t0 + nsteps) ? kBatchBlock : 1;
        if (b->flags & PFC_BATCH_NO_GRAPH) {
            if (b->prec == PFC_FP64) step_eager<double>(b, ref, len);
            else step_eager<float>(b, ref, len);
        } else if (hipGraphLaunch(b->exec[(len > 1 ? 2 : 0) + (ref ? 1 : 0)], b->s) != hipSuccess) {
            return PFC_EDEVICE;
        }
        t += len;
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(b->s) != hipSuccess)
        return PFC_EDEVICE;
    return PFC_OK;
}

int pfc_batch_shape(const pfc_batch* b, int pass, int n, int U, int* points) {
    if (!b || !points) return PFC_EINVAL;
    const auto& m = pass == 0 ? b->rclass : b->cclass;
    auto it = m.find({n, U});
    if (it == m.end()) return PFC_EINVAL;
    *points = kBatchClass[it->second].p;
    return PFC_OK;
}

}  // extern "C"
