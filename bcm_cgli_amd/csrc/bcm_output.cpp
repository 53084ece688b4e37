// Asynchronous PLOT3D output of resident cases (the reference's BCM_Output,
// main.cpp:1346-1351, writes synchronously from every rank).
//
// One output of a case costs the solver stream one layout-conversion kernel
// (SoA device state -> AoS, HBM speed).  The device-to-host copy runs on a
// second stream of default priority from a dedicated device buffer, so the
// next solver steps overlap it; a writer thread waits for the copy and writes
// the files from pinned host memory.  At most kMaxInFlight outputs are in
// flight; a further one first retires the oldest.
#include <string.h>

#include <deque>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "bcm_engine.h"

namespace bcm {

namespace {

#define HIP_TRY(x)                                                             \
  do {                                                                         \
    hipError_t _e = (x);                                                       \
    if (_e != hipSuccess) {                                                    \
      err = std::string(#x) + ": " + hipGetErrorString(_e);                    \
      return 1;                                                                \
    }                                                                          \
  } while (0)

constexpr size_t kMaxInFlight = 2;

struct PinnedBuf {
  double* p = nullptr;
  size_t bytes = 0;
};

// One queued output.  Its destructor waits for the writer thread and for any
// copy into its host buffer before releasing events and memory, so an output
// abandoned on an error path never frees memory a copy still targets.
struct OutputJob {
  std::thread th;
  PinnedBuf host;
  hipEvent_t converted = nullptr;
  hipEvent_t copied = nullptr;
  hipStream_t copy_stream = nullptr;  // set once a copy into host is enqueued
  int rc = 0;
  std::string err;
  ~OutputJob() {
    if (th.joinable()) th.join();
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    if (converted) (void)hipEventDestroy(converted);
    if (copied) (void)hipEventDestroy(copied);
    if (host.p) (void)hipHostFree(host.p);
  }
};

}  // namespace

struct OutputQueue {
  hipStream_t stream = nullptr;
  double* dev = nullptr;          // AoS staging of the case being written
  size_t dev_bytes = 0;
  hipEvent_t dev_free = nullptr;  // recorded after each copy out of dev
  bool dev_free_recorded = false;
  std::deque<std::unique_ptr<OutputJob>> jobs;
  std::vector<PinnedBuf> spare;
  std::string first_err;
};

namespace {

void retire(OutputQueue* q) {
  std::unique_ptr<OutputJob> j = std::move(q->jobs.front());
  q->jobs.pop_front();
  if (j->th.joinable()) j->th.join();
  if (j->rc && q->first_err.empty()) q->first_err = j->err;
  if (j->host.p) {
    q->spare.push_back(j->host);
    j->host = PinnedBuf();
  }
}

int take_pinned(OutputQueue* q, size_t bytes, PinnedBuf& out, std::string& err) {
  for (size_t i = 0; i < q->spare.size(); ++i)
    if (q->spare[i].bytes >= bytes) {
      out = q->spare[i];
      q->spare.erase(q->spare.begin() + i);
      return 0;
    }
  for (PinnedBuf& b : q->spare) (void)hipHostFree(b.p);
  q->spare.clear();
  out.bytes = bytes;
  HIP_TRY(hipHostMalloc((void**)&out.p, bytes, hipHostMallocDefault));
  return 0;
}

}  // namespace

int engine_output(Engine* e, int slot, int step, const std::string& dir, std::string& err) {
  HIP_TRY(hipSetDevice(e->device));
  if (slot < 0 || slot >= e->max_cases || !e->cases[slot].active) {
    err = "no such case slot";
    return 1;
  }
  if (step < 0 || step > 99999) {
    err = "output step must lie in 0..99999 (five digits in the file name)";
    return 1;
  }
  const CaseSlot& S = e->cases[slot];
  if (S.centres.size() != (size_t)3 * (S.tb.n + 1) * NP) {
    err = "the case slot holds no cell centres";
    return 1;
  }
  if (!e->out) {
    std::unique_ptr<OutputQueue> q(new OutputQueue());
    HIP_TRY(hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&q->dev_free, hipEventDisableTiming));
    e->out = q.release();
  }
  OutputQueue* q = e->out;
  while (q->jobs.size() >= kMaxInFlight) retire(q);

  const size_t n = S.tb.n;
  const size_t cube_bytes = (size_t)CUBE_STRIDE * sizeof(double);
  const size_t bytes = n * cube_bytes;
  if (q->dev_bytes < bytes) {
    HIP_TRY(hipStreamSynchronize(q->stream));
    if (q->dev) HIP_TRY(hipFree(q->dev));
    q->dev = nullptr;
    q->dev_bytes = 0;
    HIP_TRY(hipMalloc((void**)&q->dev, bytes));
    q->dev_bytes = bytes;
    q->dev_free_recorded = false;
  }

  std::unique_ptr<OutputJob> j(new OutputJob());
  HIP_TRY(hipEventCreateWithFlags(&j->converted, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&j->copied, hipEventDisableTiming));
  if (take_pinned(q, bytes + cube_bytes, j->host, err)) return 1;

  // The solver stream rewrites the staging buffer only after the previous
  // copy out of it has finished.
  if (q->dev_free_recorded) HIP_TRY(hipStreamWaitEvent(e->stream, q->dev_free, 0));
  double ref[5];
  ref_state(S.tb.ph, ref);
  if (e->prec == 1)
    launch_soa_to_aos_f(e->stream, e->Uf + (size_t)S.first * CUBE_STRIDE, q->dev, (int)n, ref);
  else
    launch_soa_to_aos(e->stream, e->U + (size_t)S.first * CUBE_STRIDE, q->dev, (int)n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(j->converted, e->stream));
  HIP_TRY(hipStreamWaitEvent(q->stream, j->converted, 0));
  HIP_TRY(hipMemcpyAsync(j->host.p + CUBE_STRIDE, q->dev, bytes, hipMemcpyDeviceToHost, q->stream));
  j->copy_stream = q->stream;
  HIP_TRY(hipEventRecord(j->copied, q->stream));
  HIP_TRY(hipEventRecord(q->dev_free, q->stream));
  q->dev_free_recorded = true;
  memset(j->host.p, 0, cube_bytes);

  OutputJob* jp = j.get();
  const int device = e->device;
  const int ncube = (int)n + 1;
  const int mpi_ncube = S.tb.mpi_ncube > 0 ? S.tb.mpi_ncube : (int)n;
  std::vector<double> centres = S.centres;
  q->jobs.push_back(std::move(j));
  jp->th = std::thread([jp, device, ncube, mpi_ncube, centres = std::move(centres), step, dir]() {
    if (hipSetDevice(device) != hipSuccess || hipEventSynchronize(jp->copied) != hipSuccess) {
      jp->rc = 1;
      jp->err = "output copy to the host failed";
      return;
    }
    const size_t blk = (size_t)ncube * NP;
    jp->rc = write_plot3d(ncube, centres.data(), centres.data() + blk, centres.data() + 2 * blk,
                          mpi_ncube, jp->host.p, step, dir, jp->err);
  });
  return 0;
}

int engine_output_wait(Engine* e, std::string& err) {
  if (!e->out) return 0;
  OutputQueue* q = e->out;
  while (!q->jobs.empty()) retire(q);
  if (!q->first_err.empty()) {
    err = q->first_err;
    q->first_err.clear();
    return 1;
  }
  return 0;
}

void engine_output_free(Engine* e) {
  if (!e->out) return;
  OutputQueue* q = e->out;
  while (!q->jobs.empty()) retire(q);
  if (q->stream) (void)hipStreamSynchronize(q->stream);
  for (PinnedBuf& b : q->spare) (void)hipHostFree(b.p);
  if (q->dev) (void)hipFree(q->dev);
  if (q->dev_free) (void)hipEventDestroy(q->dev_free);
  if (q->stream) (void)hipStreamDestroy(q->stream);
  delete q;
  e->out = nullptr;
}

}  // namespace bcm
