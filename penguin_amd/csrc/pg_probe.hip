// pg_probe.hip -- what a persistent step kernel would pay at each pass
// boundary on this GPU, measured: the grid-wide barrier of one cooperative
// launch against the gap between two kernel nodes of a captured graph.
//
// A persistent launch of the whole step would order its passes (x sweep ->
// azimuthal pass -> predictor -> corrector -> next step's prologue) with a
// grid barrier: every workgroup publishes its pass's writes to the other
// workgroups (device-scope release and acquire) and waits for all of them.
// The captured graph orders the same passes by kernel boundaries.  Both
// probes run the same pass: each thread reads a word another workgroup wrote
// in the previous pass and writes its own, so the ordering carries data, and
// the host checks the final words.  No counterpart in the reference.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/penguin_amd.h"
#include "pg_launch.h"

namespace pg {
namespace cg = cooperative_groups;

constexpr int PROBE_THREADS = 256;

// pass k: word id of buffer k&1 = (word n-1-id of buffer (k+1)&1) + 1, so
// after passes 0..K-1 buffer (K-1)&1 holds K everywhere
__device__ __forceinline__ void probe_pass(float* buf, size_t n, int k) {
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    buf[(size_t)(k & 1) * n + id] = buf[(size_t)((k + 1) & 1) * n + (n - 1 - id)] + 1.0f;
}

__global__ __launch_bounds__(PROBE_THREADS) void k_probe_grid(float* buf, int npass) {
    cg::grid_group g = cg::this_grid();
    const size_t n = (size_t)gridDim.x * blockDim.x;
    for (int k = 0; k < npass; ++k) {
        probe_pass(buf, n, k);
        g.sync();
    }
}

__global__ __launch_bounds__(PROBE_THREADS) void k_probe_pass(float* buf, int k) {
    probe_pass(buf, (size_t)gridDim.x * blockDim.x, k);
}

}  // namespace pg

namespace {

int probe_err(std::string& msg, int code, const std::string& what) {
    msg = what;
    return code;
}
int probe_fail(std::string& msg, hipError_t e, const char* what) {
    return probe_err(msg, PG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// the final words of a run of npass passes
bool probe_check(const float* dbuf, size_t n, int npass) {
    std::vector<float> h(n);
    if (hipMemcpy(h.data(), dbuf + (size_t)((npass - 1) & 1) * n, n * sizeof(float), hipMemcpyDeviceToHost) !=
        hipSuccess)
        return false;
    for (size_t i = 0; i < n; ++i)
        if (h[i] != (float)npass) return false;
    return true;
}

}  // namespace

int pg::probe_grid_barrier(int blocks, int npass, double* us_per_barrier, double* us_per_node, std::string& msg) {
    if (!us_per_barrier || !us_per_node) return probe_err(msg, PG_E_INVALID, "null argument");
    if (npass < 2 || npass > 100000)
        return probe_err(msg, PG_E_INVALID, "npass=" + std::to_string(npass) + ": use 2 .. 100000");
    int dev = 0, ncu = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_probe_grid, PROBE_THREADS, 0);
    if (e != hipSuccess) return probe_fail(msg, e, "device query");
    // a cooperative launch needs every workgroup resident at once
    if (blocks < 1 || blocks > ncu * per_cu)
        return probe_err(msg, PG_E_INVALID,
                         "blocks=" + std::to_string(blocks) + ": use 1 .. " + std::to_string(ncu * per_cu) +
                             " (co-resident workgroups of 256 threads)");
    const size_t n = (size_t)blocks * PROBE_THREADS;
    float* buf = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int rc = PG_OK;
    float ms = 0.0f;
    void* args[2] = {&buf, &npass};
    if ((e = hipMalloc(&buf, 2 * n * sizeof(float))) != hipSuccess) return probe_fail(msg, e, "hipMalloc");
    if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
        rc = probe_fail(msg, e, "stream/event");
        goto done;
    }
    // one cooperative launch of npass passes and barriers (after a warm-up)
    for (int rep = 0; rep < 2 && rc == PG_OK; ++rep) {
        if ((e = hipMemsetAsync(buf, 0, 2 * n * sizeof(float), s)) != hipSuccess ||
            (e = hipEventRecord(e0, s)) != hipSuccess ||
            (e = hipLaunchCooperativeKernel((const void*)k_probe_grid, dim3(blocks), dim3(PROBE_THREADS), args, 0, s)) !=
                hipSuccess ||
            (e = hipEventRecord(e1, s)) != hipSuccess || (e = hipStreamSynchronize(s)) != hipSuccess ||
            (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess)
            rc = probe_fail(msg, e, "cooperative launch");
        else if (!probe_check(buf, n, npass))
            rc = probe_err(msg, PG_E_STATE, "grid barrier probe: wrong words after the passes");
    }
    if (rc != PG_OK) goto done;
    *us_per_barrier = 1000.0 * ms / npass;
    // the same passes as npass kernel nodes of a captured graph
    if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
        rc = probe_fail(msg, e, "capture");
        goto done;
    }
    for (int k = 0; k < npass; ++k) hipLaunchKernelGGL(k_probe_pass, dim3(blocks), dim3(PROBE_THREADS), 0, s, buf, k);
    if ((e = hipStreamEndCapture(s, &graph)) != hipSuccess ||
        (e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0)) != hipSuccess) {
        rc = probe_fail(msg, e, "graph");
        goto done;
    }
    for (int rep = 0; rep < 2 && rc == PG_OK; ++rep) {
        if ((e = hipMemsetAsync(buf, 0, 2 * n * sizeof(float), s)) != hipSuccess ||
            (e = hipEventRecord(e0, s)) != hipSuccess || (e = hipGraphLaunch(exec, s)) != hipSuccess ||
            (e = hipEventRecord(e1, s)) != hipSuccess || (e = hipStreamSynchronize(s)) != hipSuccess ||
            (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess)
            rc = probe_fail(msg, e, "graph launch");
        else if (!probe_check(buf, n, npass))
            rc = probe_err(msg, PG_E_STATE, "graph probe: wrong words after the passes");
    }
    if (rc == PG_OK) *us_per_node = 1000.0 * ms / npass;
done:
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (s) hipStreamDestroy(s);
    hipFree(buf);
    return rc;
}
