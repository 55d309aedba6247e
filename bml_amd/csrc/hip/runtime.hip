// Device runtime for the bml HIP backend: one device, one stream (default
// priority), a device status block, and grow-on-demand scratch buffers.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdio.h>
#include <string.h>

#include "../bmlhip_abi.h"

namespace {

struct Runtime {
    int capturing = 0;
    int profiling = 0;
    int prof_n = 0;
    hipEvent_t prof_ev[2 * 4096];
    int prof_prec[4096];
    int device = -1;
    hipStream_t stream = nullptr;
    bmlhip_status *status = nullptr;
    void *scratch[BMLHIP_SCRATCH_SLOTS] = {};
    size_t scratch_bytes[BMLHIP_SCRATCH_SLOTS] = {};
    char msg[512] = {0};
    void *noalloc_ws[64] = {};  // address of `scratch` on each device
};

Runtime g;

int fail(hipError_t e, const char *what) {
    snprintf(g.msg, sizeof g.msg, "%s: %s", what, hipGetErrorString(e));
    return 3;  // BML_HIP_ERR_DEVICE
}

}  // namespace

// Workspace of the entry points that allocate nothing (for callers that
// cannot call hipMalloc: graph capture, tight serving loops).  Each GPU the
// library runs on holds its own copy until the process ends.  In host code
// the name is only a stand-in: the device address comes from
// hipGetSymbolAddress at bmlhip_init (noalloc_ws), and only that is passed
// to kernels.
__device__ __align__(256) unsigned char scratch[BMLHIP_NOALLOC_BYTES];

// Clears the per-operation fields of the status block on the library stream
// (a kernel rather than a memset node, so a captured graph replays it
// faithfully).  The overflow fields accumulate until the host has checked
// them and asks for them to be cleared too.
__global__ void k_status_reset(bmlhip_status *st, int clear_errors) {
    if (threadIdx.x == 0) {
        if (clear_errors) {
            st->m_overflow_rows = 0;
            st->first_m_overflow = INT_MAX;
        }
        st->table_overflow_rows = 0;
        st->max_nnz = 0;
        st->max_touched = 0;
        st->cursor = 0;
        st->total_nnz = 0;
        for (int k = 0; k < 4; k++) st->scalars[k] = 0.0;
    }
}

#define CHECK(call)                                   \
    do {                                              \
        hipError_t e_ = (call);                       \
        if (e_ != hipSuccess) return fail(e_, #call); \
    } while (0)

extern "C" {

const char *bmlhip_errmsg(void) { return g.msg; }

void bmlhip_set_errmsg(const char *m) { snprintf(g.msg, sizeof g.msg, "%s", m); }

int bmlhip_init(int device) {
    if (g.device >= 0) return 0;
    int n = 0;
    CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        snprintf(g.msg, sizeof g.msg, "device %d not present (%d devices)", device, n);
        return 3;
    }
    CHECK(hipSetDevice(device));
    if (device < 64 && !g.noalloc_ws[device]) {
        void *ws = nullptr;
        CHECK(hipGetSymbolAddress(&ws, HIP_SYMBOL(scratch)));
        g.noalloc_ws[device] = ws;
    }
    CHECK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    CHECK(hipMalloc((void **) &g.status, sizeof(bmlhip_status)));
    hipLaunchKernelGGL(k_status_reset, dim3(1), dim3(64), 0, g.stream, g.status, 1);
    CHECK(hipGetLastError());
    CHECK(hipStreamSynchronize(g.stream));
    g.device = device;
    return 0;
}

void bmlhip_shutdown(void) {
    if (g.device < 0) return;
    (void) hipStreamSynchronize(g.stream);
    for (int s = 0; s < BMLHIP_SCRATCH_SLOTS; s++) {
        if (g.scratch[s]) (void) hipFree(g.scratch[s]);
        g.scratch[s] = nullptr;
        g.scratch_bytes[s] = 0;
    }
    (void) hipFree(g.status);
    (void) hipStreamDestroy(g.stream);
    g.status = nullptr;
    g.stream = nullptr;
    g.device = -1;
}

void *bmlhip_stream(void) { return (void *) g.stream; }

int bmlhip_device_cus(void) {
    int n = 0;
    if (g.device < 0 ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, g.device) != hipSuccess)
        return 0;
    return n;
}

// A timed region on the library stream (the autotuner's measurements):
// begin records an event, end records a second one, waits for it and gives
// the milliseconds between them.
static hipEvent_t g_tev[2];

int bmlhip_time_begin(void) {
    if (!g_tev[0]) {
        CHECK(hipEventCreate(&g_tev[0]));
        CHECK(hipEventCreate(&g_tev[1]));
    }
    CHECK(hipEventRecord(g_tev[0], g.stream));
    return 0;
}

int bmlhip_time_end(double *ms) {
    float f = 0.0f;
    CHECK(hipEventRecord(g_tev[1], g.stream));
    CHECK(hipEventSynchronize(g_tev[1]));
    CHECK(hipEventElapsedTime(&f, g_tev[0], g_tev[1]));
    *ms = (double) f;
    return 0;
}

void *bmlhip_noalloc_workspace(size_t *bytes) {
    if (bytes) *bytes = BMLHIP_NOALLOC_BYTES;
    return g.device >= 0 && g.device < 64 ? g.noalloc_ws[g.device] : nullptr;
}

int bmlhip_malloc(void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        snprintf(g.msg, sizeof g.msg, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
        *p = nullptr;
        return 4;  // BML_HIP_ERR_ALLOC
    }
    return 0;
}

void bmlhip_free(void *p) {
    if (!p) return;
    // frees must not race kernels still reading the buffer
    (void) hipStreamSynchronize(g.stream);
    (void) hipFree(p);
}

int bmlhip_h2d(void *d, const void *s, size_t n) {
    if (!n) return 0;
    CHECK(hipMemcpyAsync(d, s, n, hipMemcpyHostToDevice, g.stream));
    CHECK(hipStreamSynchronize(g.stream));
    return 0;
}

int bmlhip_d2h(void *d, const void *s, size_t n) {
    if (!n) return 0;
    CHECK(hipMemcpyAsync(d, s, n, hipMemcpyDeviceToHost, g.stream));
    CHECK(hipStreamSynchronize(g.stream));
    return 0;
}

int bmlhip_d2d(void *d, const void *s, size_t n) {
    if (!n) return 0;
    CHECK(hipMemcpyAsync(d, s, n, hipMemcpyDeviceToDevice, g.stream));
    return 0;
}

int bmlhip_zero(void *d, size_t n) {
    if (!n) return 0;
    CHECK(hipMemsetAsync(d, 0, n, g.stream));
    return 0;
}

int bmlhip_sync(void) {
    CHECK(hipStreamSynchronize(g.stream));
    CHECK(hipGetLastError());
    return 0;
}

void *bmlhip_scratch(int slot, size_t bytes) {
    if (slot < 0 || slot >= BMLHIP_SCRATCH_SLOTS) return nullptr;
    if (bytes == 0) bytes = 256;
    if (g.scratch_bytes[slot] >= bytes) return g.scratch[slot];
    // a graph capture cannot record an allocation: the eager pass before it
    // sizes every buffer
    if (g.capturing) {
        snprintf(g.msg, sizeof g.msg, "scratch slot %d would grow during graph capture", slot);
        return nullptr;
    }
    if (g.scratch[slot]) {
        (void) hipStreamSynchronize(g.stream);
        (void) hipFree(g.scratch[slot]);
        g.scratch[slot] = nullptr;
        g.scratch_bytes[slot] = 0;
    }
    size_t want = bytes + bytes / 4;  // headroom against regrowth
    if (hipMalloc(&g.scratch[slot], want) != hipSuccess) {
        snprintf(g.msg, sizeof g.msg, "scratch allocation of %zu bytes failed", want);
        g.scratch[slot] = nullptr;
        return nullptr;
    }
    g.scratch_bytes[slot] = want;
    return g.scratch[slot];
}

void *bmlhip_scratch_zeroed(int slot, size_t bytes) {
    const bool grow = slot >= 0 && slot < BMLHIP_SCRATCH_SLOTS && g.scratch_bytes[slot] < bytes;
    void *p = bmlhip_scratch(slot, bytes);
    if (p && grow && hipMemsetAsync(p, 0, g.scratch_bytes[slot], g.stream) != hipSuccess) return nullptr;
    return p;
}

bmlhip_status *bmlhip_status_dev(void) { return g.status; }

int bmlhip_status_reset(void) {
    hipLaunchKernelGGL(k_status_reset, dim3(1), dim3(64), 0, g.stream, g.status, 0);
    CHECK(hipGetLastError());
    return 0;
}

int bmlhip_status_clear_errors(void) {
    hipLaunchKernelGGL(k_status_reset, dim3(1), dim3(64), 0, g.stream, g.status, 1);
    CHECK(hipGetLastError());
    return 0;
}

__global__ void k_prof_reset(bmlhip_status *st) {
    if (threadIdx.x < 5) {
        st->prof_products[threadIdx.x] = 0;
        st->prof_rows[threadIdx.x] = 0;
        st->prof_in_nnz[threadIdx.x] = 0;
        st->prof_out_nnz[threadIdx.x] = 0;
        st->prof_ovf_rows[threadIdx.x] = 0;
        st->prof_touched[threadIdx.x] = 0;
        st->prof_streamed[threadIdx.x] = 0;
    }
}

int bmlhip_prof_begin(void) {
    hipLaunchKernelGGL(k_prof_reset, dim3(1), dim3(64), 0, g.stream, g.status);
    CHECK(hipGetLastError());
    g.profiling = 1;
    g.prof_n = 0;
    return 0;
}

void bmlhip_prof_mark_begin(int prec) {
    if (!g.profiling || g.capturing || g.prof_n >= 4096) return;
    hipEvent_t *e = &g.prof_ev[2 * g.prof_n];
    if (hipEventCreate(&e[0]) != hipSuccess || hipEventCreate(&e[1]) != hipSuccess) return;
    g.prof_prec[g.prof_n] = prec;
    (void) hipEventRecord(e[0], g.stream);
}

void bmlhip_prof_mark_end(void) {
    if (!g.profiling || g.capturing || g.prof_n >= 4096) return;
    (void) hipEventRecord(g.prof_ev[2 * g.prof_n + 1], g.stream);
    g.prof_n++;
}

int bmlhip_prof_end(int launches[5], double kernel_ms[5]) {
    CHECK(hipStreamSynchronize(g.stream));
    for (int p = 0; p < 5; p++) {
        launches[p] = 0;
        kernel_ms[p] = 0.0;
    }
    for (int k = 0; k < g.prof_n; k++) {
        float ms = 0.0f;
        const int p = g.prof_prec[k] >= 1 && g.prof_prec[k] <= 4 ? g.prof_prec[k] : 0;
        if (hipEventElapsedTime(&ms, g.prof_ev[2 * k], g.prof_ev[2 * k + 1]) == hipSuccess) {
            kernel_ms[p] += ms;
            launches[p]++;
        }
        (void) hipEventDestroy(g.prof_ev[2 * k]);
        (void) hipEventDestroy(g.prof_ev[2 * k + 1]);
    }
    g.profiling = 0;
    g.prof_n = 0;
    return 0;
}

int bmlhip_graph_begin(void) {
    CHECK(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
    g.capturing = 1;
    return 0;
}

int bmlhip_graph_end(void **exec) {
    hipGraph_t graph = nullptr;
    g.capturing = 0;
    CHECK(hipStreamEndCapture(g.stream, &graph));
    hipGraphExec_t ge = nullptr;
    const hipError_t e = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
    (void) hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(e, "hipGraphInstantiate");
    *exec = (void *) ge;
    return 0;
}

int bmlhip_graph_launch(void *exec) {
    CHECK(hipGraphLaunch((hipGraphExec_t) exec, g.stream));
    return 0;
}

void bmlhip_graph_destroy(void *exec) {
    if (exec) (void) hipGraphExecDestroy((hipGraphExec_t) exec);
}

int bmlhip_status_fetch(bmlhip_status *h) {
    CHECK(hipMemcpyAsync(h, g.status, sizeof *h, hipMemcpyDeviceToHost, g.stream));
    CHECK(hipStreamSynchronize(g.stream));
    CHECK(hipGetLastError());
    return 0;
}

}  // extern "C"
