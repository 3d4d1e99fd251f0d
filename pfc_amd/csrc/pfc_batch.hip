// pfc_batch.hip -- the pfc_batch_* C ABI: many independent systems of uneven
// sizes stepped together, each pass one persistent launch per kernel class
// present over that class's work items (pfc_batch.hpp), each step a captured
// HIP graph, the FFT shape of every (length, U) chosen by a small autotuner.
//
// A batch borrows existing contexts: their buffers, parameters, box tables and
// NaN flags.  The contexts stay usable through the one-system calls (fields
// in and out, evaluate, box); a batch step advances all of them by the same
// steps, with the refresh of every 100th step as pfc_step does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/pfc_relax.h"
#include "pfc_batch.hpp"

extern template struct pfc::BatchLaunch<float>;
extern template struct pfc::BatchLaunch<double>;

namespace pfc {

__global__ void k_batch_tick(int* tdev) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *tdev += 1;
}

namespace {

// the two candidate classes of (length n, fields U), the default first
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

// The launches of one pass: per class present, its items' offset, count and
// persistent grid, largest classes first.
struct BatchPass {
    std::vector<int> cls, off, cnt, grid;
    BatchItem* d = nullptr;
};

struct pfc_batch {
    int n = 0, prec = 0, flags = 0;
    std::vector<pfc_ctx*> ctx;
    hipStream_t s = nullptr;
    void* dsys = nullptr;    // BatchSys<T>[n]
    void* dtune = nullptr;   // BatchSys<T>[n] with Nh, I redirected to scratch (tuning)
    BatchPass rows, cols;
    int* tdev = nullptr;
    int thost = 0;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    std::map<std::pair<int, int>, int> rclass, cclass;   // (length, U) -> class
    int cus = 256;
};

namespace {

template <typename T>
BatchSys<T> sys_of(const pfc_ctx* c, bool scratch) {
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
    b.nan = c->nan_flag;
    return b;
}

// Items of one pass for the systems in `which` (all when empty) under the class
// choice `classes`, grouped by class, each class's items permuted so that the
// blocks of one XCD (block b runs on XCD b % 8) take consecutive items.
template <typename T>
int build_pass(const pfc_batch* b, bool row_pass, const std::map<std::pair<int, int>, int>& classes,
               const std::vector<int>& which, BatchPass* out) {
    std::vector<int> sel = which;
    if (sel.empty())
        for (int i = 0; i < b->n; i++) sel.push_back(i);
    std::map<int, std::vector<BatchItem>> per;
    std::map<int, double> work;
    for (int i : sel) {
        const pfc_ctx* c = b->ctx[i];
        const int cls = classes.at({row_pass ? c->Nx : c->Ny, c->U});
        auto& v = per[cls];
        if (row_pass) {
            for (int g = 0; g < c->Ny / batch_tr(cls); g++) v.push_back({i, g});
        } else {
            for (int cb = 0; cb <= c->Nx / 2 / batch_tc(cls); cb++) v.push_back({i, cb});
        }
        work[cls] += (double)c->Nx * c->Ny * c->U;
    }
    std::vector<int> order;
    for (auto& kv : per) order.push_back(kv.first);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
    std::vector<BatchItem> all;
    BatchPass p;
    for (int cls : order) {
        std::vector<BatchItem>& v = per[cls];
        const int n8 = (int)v.size() / 8 * 8;
        if (n8 >= 8) {
            std::vector<BatchItem> q(v);
            for (int w = 0; w < n8; w++) q[w] = v[(w % 8) * (n8 / 8) + w / 8];
            v.swap(q);
        }
        int g = std::min((int)v.size(), BatchLaunch<T>::occupancy(cls, row_pass) * b->cus);
        if (g >= 8) g -= g % 8;
        p.cls.push_back(cls);
        p.off.push_back((int)all.size());
        p.cnt.push_back((int)v.size());
        p.grid.push_back(std::max(g, 1));
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
void launch_pass(const pfc_batch* b, bool row_pass, const BatchPass& p, bool ref, const void* sys) {
    for (size_t k = 0; k < p.cls.size(); k++) {
        if (row_pass)
            BatchLaunch<T>::row(p.cls[k], ref, p.d + p.off[k], p.cnt[k], (const BatchSys<T>*)sys,
                                b->tdev, p.grid[k], b->s);
        else
            BatchLaunch<T>::col(p.cls[k], ref, p.d + p.off[k], p.cnt[k], (const BatchSys<T>*)sys,
                                p.grid[k], b->s);
    }
}

// Best of 3 timed launches (after one untimed) of one pass over the systems in
// `which` with class choice `classes`, in ms; 1e30 on error.  Column passes
// run on the scratch descriptors, so no context's state changes.
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
            launch_pass<T>(b, row_pass, p, false, row_pass ? b->dsys : b->dtune);
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
    for (auto& g : groups) classes[g.first] = class_of(g.first.first, g.first.second, 0);
    for (auto& g : groups) {
        const auto key = std::make_tuple(b->prec, row_pass ? 0 : 1, g.first.first, g.first.second);
        auto hit = tuned().find(key);
        if (hit != tuned().end()) {
            classes[g.first] = hit->second;
            continue;
        }
        if (!tune) continue;
        float best = 1e30f;
        for (int k = 0; k < 2; k++) {
            std::map<std::pair<int, int>, int> trial = classes;
            trial[g.first] = class_of(g.first.first, g.first.second, k);
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
        hs[i] = sys_of<T>(b->ctx[i], false);
        ht[i] = sys_of<T>(b->ctx[i], true);
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
    int rc = build_pass<T>(b, true, b->rclass, {}, &b->rows);
    if (rc == PFC_OK) rc = build_pass<T>(b, false, b->cclass, {}, &b->cols);
    if (rc != PFC_OK) return rc;
    if (hipGetLastError() != hipSuccess) return PFC_EDEVICE;
    if (b->flags & PFC_BATCH_NO_GRAPH) return PFC_OK;
    // one graph per kind of step: the row pass, the column pass, the step counter
    for (int ref = 0; ref < 2; ref++) {
        if (hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return PFC_EDEVICE;
        launch_pass<T>(b, true, b->rows, ref != 0, b->dsys);
        launch_pass<T>(b, false, b->cols, ref != 0, b->dsys);
        hipLaunchKernelGGL(k_batch_tick, dim3(1), dim3(64), 0, b->s, b->tdev);
        if (hipStreamEndCapture(b->s, &b->graph[ref]) != hipSuccess ||
            hipGraphInstantiate(&b->exec[ref], b->graph[ref], nullptr, nullptr, 0) != hipSuccess)
            return PFC_EDEVICE;
    }
    return PFC_OK;
}

template <typename T>
void step_eager(pfc_batch* b, bool ref) {
    launch_pass<T>(b, true, b->rows, ref, b->dsys);
    launch_pass<T>(b, false, b->cols, ref, b->dsys);
    hipLaunchKernelGGL(k_batch_tick, dim3(1), dim3(64), 0, b->s, b->tdev);
}

bool batch_length(int n) { return class_of(n, 1, 0) >= 0; }

}  // namespace

extern "C" {

int pfc_batch_destroy(pfc_batch* b) {
    if (!b) return PFC_EINVAL;
    if (b->s) (void)hipStreamSynchronize(b->s);
    for (int i = 0; i < 2; i++) {
        if (b->exec[i]) (void)hipGraphExecDestroy(b->exec[i]);
        if (b->graph[i]) (void)hipGraphDestroy(b->graph[i]);
    }
    if (b->dsys) (void)hipFree(b->dsys);
    if (b->dtune) (void)hipFree(b->dtune);
    if (b->rows.d) (void)hipFree(b->rows.d);
    if (b->cols.d) (void)hipFree(b->cols.d);
    if (b->tdev) (void)hipFree(b->tdev);
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
        if (!batch_length(c->Nx) || !batch_length(c->Ny)) return PFC_ESIZE;
        if (c->sI != ilog2(c->Ny) || c->sJ != ilog2(c->Ny)) return PFC_EINVAL;
    }
    pfc_batch* b = new pfc_batch();
    b->n = n;
    b->prec = ctxs[0]->prec;
    b->flags = flags;
    b->ctx.assign(ctxs, ctxs + n);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&b->tdev, sizeof(int)) != hipSuccess) {
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
    b->thost = t0;
    if (hipMemcpyAsync(b->tdev, &b->thost, sizeof(int), hipMemcpyHostToDevice, b->s) != hipSuccess)
        return PFC_EDEVICE;
    for (int i = 0; i < nsteps; i++) {
        const bool ref = ((t0 + i) % 100 == 0);
        if (b->flags & PFC_BATCH_NO_GRAPH) {
            if (b->prec == PFC_FP64) step_eager<double>(b, ref);
            else step_eager<float>(b, ref);
        } else if (hipGraphLaunch(b->exec[ref ? 1 : 0], b->s) != hipSuccess) {
            return PFC_EDEVICE;
        }
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
