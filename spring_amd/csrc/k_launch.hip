// Launch machinery shared by the item kernels (small-cell and MFMA three-phonon kernels):
//   * persistent launch: a grid of k x (compute units) workgroups pulls 64-pair work items
//     from an atomic counter (reset by a fill kernel, so the sequence can be captured);
//   * hipGraph capture of [reset counter -> item kernel -> per-q0 reduction], replayed on
//     later calls with the same arguments;
//   * a small autotuner choosing the launch mode (one workgroup per item, or k = 1, 2, 4
//     persistent grids) and, for kernels built in several tile shapes (the s2-block size of the
//     matrix-core kernel), the shape, by timing each candidate on a sample of the problem's own
//     items; decisions persist per (kernel, Ndof, Nq, size class) in a text cache
//     (SPR_TUNE_CACHE, default $HOME/.cache/spring_amd/tune.txt).  SPR_LAUNCH_MODE and
//     SPR_TILE_SHAPE force a choice.
// Every per-item partial lands at its item index and the reduction sums items in index
// order, so the result does not depend on which workgroup took which item.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <algorithm>
#include <map>
#include <sstream>
#include <string>
#include <sys/stat.h>

#include "spr_internal.h"

namespace spr {

__global__ void k_fill_int(int *p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

static std::string tune_path() {
    const char *e = getenv("SPR_TUNE_CACHE");
    if (e && *e) return e;
    const char *h = getenv("HOME");
    return std::string(h ? h : "/tmp") + "/.cache/spring_amd/tune.txt";
}

static std::map<std::string, int> &tune_table() {
    static std::map<std::string, int> t;
    static bool loaded = false;
    if (!loaded) {
        loaded = true;
        std::ifstream f(tune_path());
        std::string k;
        int v;
        while (f >> k >> v) t[k] = v;
    }
    return t;
}

static void tune_store(const std::string &k, int v) {
    tune_table()[k] = v;
    const std::string p = tune_path();
    const size_t slash = p.find_last_of('/');
    if (slash != std::string::npos) {
        std::string dir = p.substr(0, slash);
        // create parent directories (two levels are enough for the default path)
        const size_t s2 = dir.find_last_of('/');
        if (s2 != std::string::npos) mkdir(dir.substr(0, s2).c_str(), 0755);
        mkdir(dir.c_str(), 0755);
    }
    std::ofstream f(p, std::ios::trunc);
    for (auto &kv : tune_table()) f << kv.first << " " << kv.second << "\n";
}

static int grid_for(Ctx &c, const ItemLaunch &L, int mode) {
    if (mode == 0) return L.nitems;
    const long g = (long)mode * c.ncu * std::max(1, L.per_cu);
    return (int)std::min<long>(g, L.nitems);
}

static int enqueue(Ctx &c, ItemLaunch &L, int mode, int q0_first, int q0_count, int it0, double *d_lw,
                   bool timed) {
    ItemCommon *ic = reinterpret_cast<ItemCommon *>(L.args);
    ic->nitems = L.nitems;
    ic->counter = mode == 0 ? nullptr : c.counter;
    if (mode) hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c.stream, c.counter, 1, 0);
    if (timed) SPR_HIP(hipEventRecord(c.ev0, c.stream));
    void *args[5] = {L.args, L.extra[0], L.extra[1], L.extra[2], L.extra[3]};
    SPR_HIP(hipLaunchKernel(L.func, dim3(grid_for(c, L, mode)), dim3(L.block), args, L.dyn, c.stream));
    if (timed) SPR_HIP(hipEventRecord(c.ev1, c.stream));
    if (d_lw) return launch_reduce_items(c, q0_first, q0_count, it0, L.upi, d_lw);
    return SPR_OK;
}

// Times one launch mode on the first `sub` items (enough to show the persistent grid's
// effect without running the whole problem once per candidate).
static int time_mode(Ctx &c, ItemLaunch &L, int mode, int sub, float *ms) {
    int rc;
    const int full = L.nitems;
    L.nitems = sub;
    rc = enqueue(c, L, mode, 0, 0, 0, nullptr, true);
    L.nitems = full;
    if (rc) return rc;
    SPR_HIP(hipEventSynchronize(c.ev1));
    SPR_HIP(hipEventElapsedTime(ms, c.ev0, c.ev1));
    return SPR_OK;
}

static std::string size_key(const Ctx &c, const std::string &name, int nitems) {
    // the decision is kept per size class (item counts within a factor of two share it), so
    // equal-sized slices of one wedge are tuned once
    int size_class = 0;
    while ((2 << size_class) <= nitems) ++size_class;
    return name + "|nd" + std::to_string(c.nd) + "|nq" + std::to_string(c.nq) + "|items2^" +
           std::to_string(size_class);
}

int run_items_shapes(Ctx &c, ItemLaunch *Ls, int n, int q0_first, int q0_count, int it0, double *d_lw) {
    if (n == 1 || Ls[0].direct) return run_items(c, Ls[0], q0_first, q0_count, it0, d_lw);
    int rc;
    const int items = Ls[0].nitems / Ls[0].upi;
    const std::string key = size_key(c, Ls[0].name + "|shape", items);
    const char *fs = getenv("SPR_TILE_SHAPE");
    const char *fm = getenv("SPR_LAUNCH_MODE");
    const char *at = getenv("SPR_AUTOTUNE");
    int shape = 0, mode = -1;
    if (fs && *fs) {
        shape = std::min(std::max(atoi(fs), 0), n - 1);
    } else if (tune_table().count(key)) {
        const int v = tune_table()[key];
        shape = std::min(v / 8, n - 1);
        mode = v % 8;
    } else if (!(at && std::string(at) == "0") && items >= 4 * c.ncu / std::max(1, Ls[0].upi)) {
        // every candidate runs the same first `sub` items (its own units of them)
        const int sub = std::min(items, 4 * c.ncu);
        float best = 1e30f, ms;
        int best_v = 0;
        for (int v = 0; v < n; ++v) {
            ItemLaunch &L = Ls[v];
            if (L.dyn > 64 * 1024)
                SPR_HIP(hipFuncSetAttribute(L.func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.dyn));
            if ((rc = time_mode(c, L, 2, sub * L.upi, &ms))) return rc;   // warm-up (code load, caches)
            for (int m : {0, 1, 2, 4}) {
                if ((rc = time_mode(c, L, m, sub * L.upi, &ms))) return rc;
                if (ms < best) {
                    best = ms;
                    best_v = v * 8 + m;
                }
            }
        }
        shape = best_v / 8;
        mode = best_v % 8;
        tune_store(key, best_v);
    }
    if (fm && *fm) mode = -1;   // run_items applies SPR_LAUNCH_MODE itself
    return run_items(c, Ls[shape], q0_first, q0_count, it0, d_lw, mode);
}

int run_items(Ctx &c, ItemLaunch &L, int q0_first, int q0_count, int it0, double *d_lw, int preset) {
    int rc;
    c.last_upi = L.upi;
    c.last_shape = L.shape;
    c.last_name = L.name;
    if (L.dyn > 64 * 1024)
        SPR_HIP(hipFuncSetAttribute(L.func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.dyn));
    // ---- launch mode: env override, tuned value, or default (persistent 2 x CUs) ----
    if (L.direct) {
        if ((rc = enqueue(c, L, 0, 0, 0, 0, nullptr, true))) return rc;
        SPR_HIP(hipEventSynchronize(c.ev1));
        return SPR_OK;
    }
    int mode = 2;
    const std::string key = size_key(c, L.name + L.shape, L.nitems);
    const char *force = getenv("SPR_LAUNCH_MODE");
    const char *at = getenv("SPR_AUTOTUNE");
    if (force && *force) {
        mode = atoi(force);
        if (mode != 0 && mode != 1 && mode != 2 && mode != 4) mode = 2;
    } else if (preset >= 0) {
        mode = preset;
    } else if (tune_table().count(key)) {
        mode = tune_table()[key];
    } else if (!(at && std::string(at) == "0") && L.nitems >= 4 * c.ncu) {
        float best = 1e30f;
        int best_mode = 2;
        float ms;
        const int sub = std::min(L.nitems, 8 * c.ncu * std::max(1, L.per_cu));
        if ((rc = time_mode(c, L, 2, sub, &ms))) return rc;   // warm-up (code load, caches)
        for (int m : {0, 1, 2, 4}) {
            if ((rc = time_mode(c, L, m, sub, &ms))) return rc;
            if (ms < best) {
                best = ms;
                best_mode = m;
            }
        }
        mode = best_mode;
        tune_store(key, mode);
    }
    c.last_mode = mode;
    // ---- graph replay (same kernel, arguments and mode) or direct launch ----
    const char *ng = getenv("SPR_NO_GRAPH");
    const bool use_graph = !(ng && std::string(ng) == "1");
    ItemCommon *ic = reinterpret_cast<ItemCommon *>(L.args);
    ic->nitems = L.nitems;
    ic->counter = mode == 0 ? nullptr : c.counter;
    std::string gkey;
    if (use_graph) {
        std::ostringstream os;
        os << key << "|m" << mode << "|q" << q0_first << ":" << q0_count << "|out" << (const void *)d_lw << "|";
        const unsigned char *b = reinterpret_cast<const unsigned char *>(L.args);
        for (size_t i = 0; i < L.args_size; ++i) os << std::hex << (int)b[i] << ".";
        for (int e = 0; e < L.nextra; ++e) os << "|" << *reinterpret_cast<void **>(L.extra[e]);
        gkey = os.str();
    }
    if (use_graph) {
        SPR_HIP(hipEventRecord(c.ev0, c.stream));
        hipGraphExec_t exec = nullptr;
        for (auto &g : c.graphs)
            if (g.first == gkey) exec = g.second;
        if (!exec) {
            hipGraph_t graph;
            SPR_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue(c, L, mode, q0_first, q0_count, it0, d_lw, false);
            hipError_t e = hipStreamEndCapture(c.stream, &graph);
            if (rc) return rc;
            if (e != hipSuccess) {
                set_error(std::string("graph capture failed: ") + hipGetErrorString(e));
                return SPR_E_HIP;
            }
            SPR_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            SPR_HIP(hipGraphDestroy(graph));
            c.graphs.emplace_back(gkey, exec);
        }
        SPR_HIP(hipGraphLaunch(exec, c.stream));
        SPR_HIP(hipEventRecord(c.ev1, c.stream));
    } else {
        // direct launch: the events bracket the item kernel alone
        if ((rc = enqueue(c, L, mode, q0_first, q0_count, it0, d_lw, true))) return rc;
    }
    if (c.defer_timing) {
        // the caller synchronises later and then reads the events (collect_timing)
        c.timing_pending = true;
        return SPR_OK;
    }
    SPR_HIP(hipEventSynchronize(c.ev1));
    float ms = 0.f;
    SPR_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    c.launches += 1;
    c.kernel_ms += ms;
    return SPR_OK;
}

int collect_timing(Ctx &c) {
    if (!c.timing_pending) return SPR_OK;
    c.timing_pending = false;
    float ms = 0.f;
    SPR_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    c.launches += 1;
    c.kernel_ms += ms;
    return SPR_OK;
}

void drop_graphs(Ctx &c) {
    for (auto &g : c.graphs) (void)hipGraphExecDestroy(g.second);
    c.graphs.clear();
}

}  // namespace spr
