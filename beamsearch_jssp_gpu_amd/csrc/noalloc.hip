// noalloc.hip -- jssp_solve_noalloc: a single device solve that allocates
// nothing.  Its workspace (beams, keys, records, histograms, trail, the
// instance descriptor and the op table) is carved from a file-scope static
// device array of fixed size instead of the engine's growable arena, the
// instance goes up by plain copies from the caller's arrays, and the results
// come back into the caller's result struct: no hipMalloc, hipHostMalloc,
// hipFree or any other allocating call.
//
// The array is per device and per process (each process loads its own copy
// of the code object) and lives until the process ends.  Its device address
// is taken once per device, with that device current, by
// noalloc_setup (called from jssp_engine_create); host code only ever uses
// that address, never the array's name (which on the host is a stand-in,
// not a device address).  Calls on one device share the array, so this file
// serialises them with a per-device lock held for the whole (synchronous)
// call.  Nothing in the array needs clearing between calls: the init kernel
// of every solve resets the counters, histograms and selection state it
// uses, as it does for the reused engine arena.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "engine.hpp"
#include "level_kernels.hpp"
#include "workspace.hpp"

namespace {

constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[JSSP_NOALLOC_SCRATCH_BYTES];

struct DeviceScratch {
  std::mutex mu;        // serialises calls on the device (they share the array)
  bool set = false;     // noalloc_setup ran for the device
  void* ptr = nullptr;  // device address of `scratch` on the device
  hipError_t err = hipSuccess;
};
DeviceScratch g_dev[kMaxDevices];
std::mutex g_setup_mu;

// (machine, duration) pairs of the op table from the two uploaded arrays.
__global__ __launch_bounds__(256) void k_pack_ops(const int* mach, const int* dur, int2* ops, int total) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) ops[i] = make_int2(mach[i], dur[i]);
}

}  // namespace

namespace jssp {

// Takes the device address of `scratch` on `device`, which must be current.
// Once per device; later calls return the recorded result.
hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_setup_mu);
  DeviceScratch& d = g_dev[device];
  if (!d.set) {
    void* ptr = nullptr;
    d.err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    d.ptr = (d.err == hipSuccess) ? ptr : nullptr;
    if (d.err != hipSuccess) (void)hipGetLastError();
    d.set = true;
  }
  return d.err;
}

}  // namespace jssp

extern "C" int jssp_solve_noalloc(jssp_engine* e, const jssp_instance* in, const jssp_params* p,
                                  jssp_result* res) {
  using namespace jssp;
  if (!e) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> elk(e->mu);
  int rc = check_instance(e, in);
  if (rc) return rc;
  rc = check_params(e, in, p);
  if (rc) return rc;
  if (!res) return set_error(e, JSSP_ERR_INVALID, "jssp_solve_noalloc", "null result");
  if (p->mode != JSSP_MODE_ICH_DET && p->mode != JSSP_MODE_PMS_GPU)
    return set_error(e, JSSP_ERR_INVALID, "jssp_solve_noalloc", "modes ICH_DET and PMS_GPU only");
  if (p->n_seeds > 1 || p->n_passes > 1)
    return set_error(e, JSSP_ERR_INVALID, "jssp_solve_noalloc", "one beam (n_seeds, n_passes <= 1) only");
  if (e->device < 0 || e->device >= kMaxDevices)
    return set_error(e, JSSP_ERR_INVALID, "jssp_solve_noalloc", "device index out of range");
  DeviceScratch& ds = g_dev[e->device];
  if (!ds.set || !ds.ptr)
    return set_error(e, JSSP_ERR_HIP, "jssp_solve_noalloc",
                     std::string("scratch array address unavailable: ") +
                         (ds.set ? hipGetErrorString(ds.err) : "noalloc_setup did not run"));

  // The plan: the solve's workspace, then the descriptor and the two
  // instance arrays, each 256-byte aligned.
  const Shape s = make_shape(in, p, 0);
  const int total = s.n * s.m;
  Work w;
  Bump b;
  b.off = carve(s, nullptr, w);
  b.take<InstDesc>(1);
  b.take<int>(total);
  b.take<int>(total);
  const size_t need = b.off + 256;
  if (need > (size_t)JSSP_NOALLOC_SCRATCH_BYTES)
    return set_error(e, JSSP_ERR_LIMIT, "jssp_solve_noalloc",
                     "workspace of " + std::to_string(need) + " bytes exceeds the " +
                         std::to_string((long long)JSSP_NOALLOC_SCRATCH_BYTES) +
                         "-byte scratch array (use jssp_solve, or a smaller beam / the two-phase path)");

  std::lock_guard<std::mutex> dlk(ds.mu);
  JSSP_HIP(e, hipSetDevice(e->device));
  char* base = static_cast<char*>(ds.ptr);
  Bump t;
  t.base = base;
  t.off = carve(s, base, w);
  InstDesc* d_desc = t.take<InstDesc>(1);
  int* d_mach = t.take<int>(total);
  int* d_dur = t.take<int>(total);

  // Instance and descriptor up by plain copies from the caller's memory and
  // this frame; the call waits for its stream before returning, on every
  // path, so both stay valid while the copies run.
  InstDesc desc = make_desc(s, w, p, in);
  desc.base.jump = w.jump;
  int h_res[4] = {0, 0, 0, 0};
  unsigned long long h_exp = 0;
  auto enqueue = [&]() -> int {
    JSSP_HIP(e, hipMemcpyAsync(d_mach, in->machine, (size_t)total * sizeof(int), hipMemcpyHostToDevice, e->stream));
    JSSP_HIP(e, hipMemcpyAsync(d_dur, in->duration, (size_t)total * sizeof(int), hipMemcpyHostToDevice, e->stream));
    k_pack_ops<<<(total + 255) / 256, 256, 0, e->stream>>>(d_mach, d_dur, w.ops, total);
    JSSP_HIP(e, hipGetLastError());
    JSSP_HIP(e, hipMemcpyAsync(d_desc, &desc, sizeof(InstDesc), hipMemcpyHostToDevice, e->stream));
    const int r = enqueue_single_solve(e, s, p, d_desc, nullptr);
    if (r) return r;
    JSSP_HIP(e, hipMemcpyAsync(h_res, w.result, 3 * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    JSSP_HIP(e, hipMemcpyAsync(&h_exp, w.expanded, sizeof(h_exp), hipMemcpyDeviceToHost, e->stream));
    return JSSP_OK;
  };
  rc = enqueue();
  const hipError_t sync = hipStreamSynchronize(e->stream);
  if (rc) return rc;
  if (sync != hipSuccess) return set_error(e, JSSP_ERR_HIP, "hipStreamSynchronize", hipGetErrorString(sync));
  if (h_res[1] && res->sequence)
    JSSP_HIP(e, hipMemcpy(res->sequence, w.seq, (size_t)s.L * sizeof(int), hipMemcpyDeviceToHost));
  float msec = 0.f;
  JSSP_HIP(e, hipEventElapsedTime(&msec, e->ev_start, e->ev_stop));
  res->makespan = h_res[0];
  res->found = h_res[1];
  res->expanded_states = (int64_t)h_exp;
  res->levels = s.L;
  res->best_seed_index = 0;
  res->device_ms = msec;
  return JSSP_OK;
}
