// A HIP runtime for the host side of libautoplace_gp alone: device memory is host memory, kernels do not run, and
// streams, events, graphs and their executables are heap objects whose lifetimes the stub checks. Built with the
// library's sources compiled for the host only (hipcc --cuda-host-only) and AddressSanitizer, it runs the level's
// bookkeeping -- graph cache, capture, events, streams, buffers -- on the CPU, where ASan sees a use after free or a
// double free, and the stub's own checks see:
//   * an executable graph destroyed while a launch of it has not been waited for (a synchronisation of its stream,
//     of a stream that waited on an event recorded after it, of the device, or of such an event);
//   * a stream, event, graph or executable used after its destruction (a magic word, and ASan: it is freed);
//   * a capture ended on a stream that did not begin it, or work captured into a graph already ended.
// Every check failure prints what it found and aborts.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <set>
#include <vector>

namespace {

std::mutex mu;

[[noreturn]] void fail(const char* what) {
    fprintf(stderr, "hip stub: %s\n", what);
    fflush(stderr);
    abort();
}

struct Graph;
struct Exec;

// one launch of an executable graph; done once a synchronisation that covers it has returned
struct Launch {
    Exec* exec;
    bool done = false;
};
using LaunchRef = std::shared_ptr<Launch>;

struct Stream {
    unsigned magic = 0x5743;
    Graph* capturing = nullptr;            // the graph being captured on this stream (null: none)
    bool origin = false;                   // began the capture
    std::vector<LaunchRef> pending;        // launches enqueued on it (or waited for through an event) since its last sync
    void touch() const {
        if (magic != 0x5743) fail("stream used after destruction");
    }
};

struct Event {
    unsigned magic = 0x4576;
    Graph* captured_in = nullptr;          // recorded inside a capture: waiting on it joins the capture
    std::vector<LaunchRef> covers;         // the launches before its record on that stream
    void touch() const {
        if (magic != 0x4576) fail("event used after destruction");
    }
};

struct Graph {
    unsigned magic = 0x4772;
    bool ended = false;
    long nodes = 0;
    std::vector<Event*> events;            // event-record nodes
    void touch() const {
        if (magic != 0x4772) fail("graph used after destruction");
    }
};

struct Exec {
    unsigned magic = 0x4578;
    std::vector<LaunchRef> launches;       // its launches not yet known done
    std::vector<Event*> events;            // the events its record nodes record
    void touch() const {
        if (magic != 0x4578) fail("executable graph used after destruction");
    }
};

std::set<Stream*> streams;
Stream null_stream;

Stream* S(hipStream_t s) {
    Stream* p = s ? (Stream*)s : &null_stream;
    p->touch();
    return p;
}

void drain(Stream* s) {
    for (LaunchRef& l : s->pending) l->done = true;
    s->pending.clear();
}

thread_local hipError_t last = hipSuccess;
struct Config {
    dim3 grid, block;
    size_t shmem;
    hipStream_t stream;
};
thread_local std::vector<Config> configs;
alignas(256) unsigned char symbol_storage[1 << 20];

}  // namespace

extern "C" {

// ---------------------------------------------------------------- registration and launches (no device code)
void** __hipRegisterFatBinary(const void*) {
    static void* handle = nullptr;
    return &handle;
}
void __hipUnregisterFatBinary(void**) {}
void __hipRegisterFunction(void**, const void*, char*, const char*, unsigned, void*, void*, void*, void*, int*) {}
void __hipRegisterVar(void**, void*, char*, const char*, int, size_t, int, int) {}
hipError_t __hipPushCallConfiguration(dim3 grid, dim3 block, size_t shmem, hipStream_t stream) {
    configs.push_back(Config{grid, block, shmem, stream});
    return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3* grid, dim3* block, size_t* shmem, hipStream_t* stream) {
    if (configs.empty()) fail("launch configuration popped without a push");
    const Config c = configs.back();
    configs.pop_back();
    *grid = c.grid;
    *block = c.block;
    *shmem = c.shmem;
    *stream = c.stream;
    return hipSuccess;
}
hipError_t hipLaunchKernel(const void*, dim3 grid, dim3 block, void**, size_t, hipStream_t stream) {
    std::lock_guard<std::mutex> g(mu);
    Stream* s = S(stream);
    if (grid.x == 0 || block.x == 0 || block.x > 1024) fail("kernel launched with an empty or oversized grid");
    if (s->capturing) {
        s->capturing->touch();
        if (s->capturing->ended) fail("kernel captured into an ended graph");
        s->capturing->nodes++;
    }
    return hipSuccess;
}
hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int value) {
    if (value > 160 * 1024) fail("dynamic LDS above 160 KiB asked for");
    return hipSuccess;
}
hipError_t hipGetLastError(void) {
    const hipError_t e = last;
    last = hipSuccess;
    return e;
}
const char* hipGetErrorString(hipError_t) { return "hip stub error"; }

// ---------------------------------------------------------------- device
hipError_t hipGetDevice(int* d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidDevice; }
int hipGetStreamDeviceId(hipStream_t) { return 0; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
    switch (a) {
        case hipDeviceAttributeWarpSize: *v = 64; break;
        case hipDeviceAttributeMaxSharedMemoryPerBlock: *v = 65536; break;
        case hipDeviceAttributeMultiprocessorCount: *v = 256; break;
        case hipDeviceAttributeMaxThreadsPerBlock: *v = 1024; break;
        default: *v = 1; break;
    }
    return hipSuccess;
}
hipError_t hipGetDevicePropertiesR0600(hipDeviceProp_tR0600* p, int) {
    std::memset(p, 0, sizeof(*p));
    std::strcpy(p->name, "host stub");
    std::strcpy(p->gcnArchName, "gfx950");
    p->warpSize = 64;
    p->multiProcessorCount = 256;
    p->maxThreadsPerBlock = 1024;
    p->sharedMemPerBlock = 65536;
    p->totalGlobalMem = (size_t)1 << 34;
    return hipSuccess;
}
hipError_t hipDeviceSynchronize(void) {
    std::lock_guard<std::mutex> g(mu);
    drain(&null_stream);
    for (Stream* s : streams) drain(s);
    return hipSuccess;
}

// ---------------------------------------------------------------- memory (host memory; kernels never touch it)
hipError_t hipMalloc(void** p, size_t n) {
    *p = std::calloc(1, n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
    hipDeviceSynchronize();
    std::free(p);
    return hipSuccess;
}
hipError_t hipMallocAsync(void** p, size_t n, hipStream_t s) {
    S(s);
    return hipMalloc(p, n);
}
hipError_t hipFreeAsync(void* p, hipStream_t s) {
    S(s);
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned int) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) {
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
    if (n) std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st) {
    S(st);
    return hipMemcpy(d, s, n, k);
}
hipError_t hipMemset(void* d, int v, size_t n) {
    if (n) std::memset(d, v, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t s) {
    Stream* p = S(s);
    if (p->capturing) fail("memset captured into a graph");
    return hipMemset(d, v, n);
}
hipError_t hipGetSymbolAddress(void** p, const void*) {
    *p = symbol_storage;
    return hipSuccess;
}

// ---------------------------------------------------------------- streams
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
    std::lock_guard<std::mutex> g(mu);
    Stream* p = new Stream();
    streams.insert(p);
    *s = (hipStream_t)p;
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    std::lock_guard<std::mutex> g(mu);
    Stream* p = S(s);
    if (p->capturing) fail("stream destroyed while capturing");
    drain(p);
    streams.erase(p);
    p->magic = 0;
    delete p;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
    std::lock_guard<std::mutex> g(mu);
    Stream* p = S(s);
    if (p->capturing) fail("stream synchronised while capturing");
    drain(p);
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned int) {
    std::lock_guard<std::mutex> g(mu);
    Stream* p = S(s);
    Event* ev = (Event*)e;
    ev->touch();
    if (ev->captured_in && !ev->captured_in->ended && !p->capturing) {
        p->capturing = ev->captured_in;               // joins the capture
        p->origin = false;
    }
    for (LaunchRef& l : ev->covers) p->pending.push_back(l);   // a sync of p now also covers them
    return hipSuccess;
}

// ---------------------------------------------------------------- events
hipError_t hipEventCreate(hipEvent_t* e) {
    *e = (hipEvent_t) new Event();
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned int) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
    std::lock_guard<std::mutex> g(mu);
    Event* ev = (Event*)e;
    ev->touch();
    ev->magic = 0;
    delete ev;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    std::lock_guard<std::mutex> g(mu);
    Event* ev = (Event*)e;
    ev->touch();
    Stream* p = S(s);
    ev->captured_in = p->capturing;
    ev->covers = p->capturing ? std::vector<LaunchRef>() : p->pending;
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
    std::lock_guard<std::mutex> g(mu);
    Event* ev = (Event*)e;
    ev->touch();
    for (LaunchRef& l : ev->covers) l->done = true;
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    ((Event*)a)->touch();
    ((Event*)b)->touch();
    *ms = 0.125f;
    return hipSuccess;
}

// ---------------------------------------------------------------- capture and graphs
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    std::lock_guard<std::mutex> g(mu);
    Stream* p = S(s);
    if (p->capturing) fail("capture begun on a capturing stream");
    p->capturing = new Graph();
    p->origin = true;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* out) {
    std::lock_guard<std::mutex> g(mu);
    Stream* p = S(s);
    if (!p->capturing || !p->origin) fail("capture ended on a stream that did not begin it");
    Graph* gr = p->capturing;
    gr->ended = true;
    for (Stream* q : streams)
        if (q->capturing == gr) q->capturing = nullptr;
    if (null_stream.capturing == gr) null_stream.capturing = nullptr;
    p->capturing = nullptr;
    p->origin = false;
    *out = (hipGraph_t)gr;
    return hipSuccess;
}
hipError_t hipStreamGetCaptureInfo_v2(hipStream_t s, hipStreamCaptureStatus* st, unsigned long long* id,
                                      hipGraph_t* graph, const hipGraphNode_t** deps, size_t* nd) {
    std::lock_guard<std::mutex> g(mu);
    Stream* p = S(s);
    *st = p->capturing ? hipStreamCaptureStatusActive : hipStreamCaptureStatusNone;
    if (id) *id = (unsigned long long)(uintptr_t)p->capturing;
    if (graph) *graph = (hipGraph_t)p->capturing;
    if (deps) *deps = nullptr;
    if (nd) *nd = 0;
    return hipSuccess;
}
hipError_t hipGraphAddEventRecordNode(hipGraphNode_t* node, hipGraph_t graph, const hipGraphNode_t*, size_t,
                                      hipEvent_t e) {
    std::lock_guard<std::mutex> g(mu);
    Graph* gr = (Graph*)graph;
    gr->touch();
    ((Event*)e)->touch();
    gr->events.push_back((Event*)e);
    gr->nodes++;
    *node = (hipGraphNode_t)(uintptr_t)gr->nodes;
    return hipSuccess;
}
hipError_t hipStreamUpdateCaptureDependencies(hipStream_t s, hipGraphNode_t*, size_t, unsigned int) {
    S(s);
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* out, hipGraph_t graph, hipGraphNode_t*, char*, size_t) {
    std::lock_guard<std::mutex> g(mu);
    Graph* gr = (Graph*)graph;
    gr->touch();
    if (!gr->ended) fail("graph instantiated before its capture ended");
    Exec* x = new Exec();
    x->events = gr->events;
    *out = (hipGraphExec_t)x;
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t s) {
    std::lock_guard<std::mutex> g(mu);
    Exec* x = (Exec*)exec;
    x->touch();
    for (Event* e : x->events) e->touch();            // its record nodes record these events
    Stream* p = S(s);
    LaunchRef l = std::make_shared<Launch>();
    l->exec = x;
    x->launches.push_back(l);
    p->pending.push_back(l);
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t exec) {
    std::lock_guard<std::mutex> g(mu);
    Exec* x = (Exec*)exec;
    x->touch();
    for (const LaunchRef& l : x->launches)
        if (!l->done) fail("executable graph destroyed while a launch of it may still run");
    x->magic = 0;
    delete x;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t graph) {
    std::lock_guard<std::mutex> g(mu);
    Graph* gr = (Graph*)graph;
    gr->touch();
    gr->magic = 0;
    delete gr;
    return hipSuccess;
}

}  // extern "C"
