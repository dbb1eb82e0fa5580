// sk_tune.hip — the execution-plan autotuner (sk_hip_autotune): times the scan of one workload under each candidate
// plan on this device and reports the fastest. Built on the public C ABI only; no counterpart in the reference, whose
// Metal engine sizes its batches from a fixed byte budget (MetalEngine.swift:434-455).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "sk_internal.h"

namespace sk {
namespace {

struct DeviceOutputs {
  std::vector<void*> ptrs;
  ~DeviceOutputs() {
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
  }
};

// positions spread evenly over the extent (row-major, x slowest, as GridScan orders them)
std::vector<double> spread_positions(const sk_grid& g, int npos) {
  const int side = std::max(1, (int)std::ceil(std::sqrt((double)npos)));
  std::vector<double> p(2 * (size_t)npos);
  for (int k = 0; k < npos; ++k) {
    p[2 * k] = ((k / side) + 0.5) * g.extent_x / side;
    p[2 * k + 1] = ((k % side) + 0.5) * g.extent_y / side;
  }
  return p;
}

double scan_seconds(const sk_potential* const* pots, int npots, const sk_probe* probe, int npos, const double* pos,
                    const sk_detector* dets, int ndets, void* const* outs, const sk_options& o, int reps) {
  auto run = [&] {
    const int rc = sk_hip_scan(pots, npots, probe, npos, pos, dets, ndets, outs, 1, &o);
    if (rc != SK_OK) fail(rc, std::string("autotune: a candidate scan failed: ") + last_error());
    SK_HIP(hipDeviceSynchronize());
  };
  run();  // warm: workspaces, plans, graph capture
  const auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < reps; ++r) run();
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count() / reps;
}

// "<architecture> <compute units> CU" of the library's device (device 0, as the context selects)
void describe_device(char* out, size_t len) {
  hipDeviceProp_t prop;
  SK_HIP(hipGetDeviceProperties(&prop, 0));
  std::string arch = prop.gcnArchName;
  arch = arch.substr(0, arch.find(':'));  // drop the feature suffix (":sramecc+:xnack-")
  std::snprintf(out, len, "%s %d CU", arch.c_str(), prop.multiProcessorCount);
}

}  // namespace
}  // namespace sk

extern "C" int sk_hip_autotune(const sk_potential* const* pots, int npots, const sk_probe* probe,
                               const sk_detector* dets, int ndets, int npos, sk_tune_report* report) {
  using namespace sk;
  return guarded_call([&] {
    if (!report) fail(SK_ERR_INVALID, "autotune: no report");
    if (!pots || npots < 1 || !pots[0]) fail(SK_ERR_INVALID, "autotune: no potential");
    if (!probe) fail(SK_ERR_INVALID, "autotune: no probe");
    if (ndets < 0 || (ndets > 0 && !dets)) fail(SK_ERR_INVALID, "autotune: invalid detectors");
    if (sk_hip_available() != 1) fail(SK_ERR_UNAVAILABLE, std::string("autotune: ") + sk_hip_last_error());
    std::memset(report, 0, sizeof(*report));
    if (npos <= 0) npos = 4096;
    sk_grid g{};
    if (sk_hip_potential_grid(pots[0], &g) != SK_OK) fail(SK_ERR_INVALID, "autotune: invalid potential");
    long long slices = 0;
    for (int k = 0; k < npots; ++k) {
      if (!pots[k]) fail(SK_ERR_INVALID, "autotune: null potential");
      slices += sk_hip_potential_slice_count(pots[k]);
    }
    // detector outputs on the device, sized as the scan writes them
    DeviceOutputs outs;
    outs.ptrs.assign(std::max(1, ndets), nullptr);
    for (int d = 0; d < ndets; ++d) {
      int shape[2];
      const int count = sk_hip_detector_shape(&g, probe->energy, &dets[d], shape);
      if (count < 0) fail(count, std::string("autotune: ") + last_error());
      SK_HIP(hipMalloc(&outs.ptrs[d], sizeof(double) * (size_t)npos * std::max(1, count)));
    }
    const std::vector<double> pos = spread_positions(g, npos);
    // candidates: positions per launch x column tile width (512-point columns) x graph replay x lane pipeline x streams
    size_t free_b = 0, total_b = 0;
    SK_HIP(hipMemGetInfo(&free_b, &total_b));
    const double wave_bytes = 8.0 * g.nx * g.ny;
    // the library's default (batch 0: sized to stay in the last-level cache, split over two streams where the scan
    // can) and fixed sizes that stream from HBM
    std::vector<int> batches{0};
    const int cached = default_scan_batch(*pots[0], npos);
    for (int b : {512, 1024, 2048, 4096})
      if (b <= std::max(512, npos) && b != cached && 3.0 * b * wave_bytes < 0.5 * (double)free_b)
        batches.push_back(b);
    std::vector<int> tiles = g.nx == 512 ? std::vector<int>{8, 16} : std::vector<int>{0};
    std::vector<int> pipes = npots > 1 ? std::vector<int>{0, 1} : std::vector<int>{0};
    // the library's default batch also runs on one stream (streams = 1) beside its own choice of streams
    std::vector<sk_options> cand;
    for (int b : batches)
      for (int t : tiles)
        for (int gr : {0, 1})
          for (int pp : pipes)
            for (int st : {0, 1}) {
              if (st == 1 && (b != 0 || pp != 0)) continue;
              sk_options o{};
              o.batch = b;
              o.tile = b;  // a fixed batch runs whole per kernel launch, streaming from HBM
              o.use_graph = gr;
              o.fast_path = 1;
              o.pipeline = pp;
              o.col_tile = t;
              o.streams = st;
              if ((int)cand.size() < SK_TUNE_MAX) cand.push_back(o);
            }
    const int reps = 2;
    int best = -1;
    for (size_t i = 0; i < cand.size(); ++i) {
      sk_options o = cand[i];
      double secs;
      // a lane pipeline the grid cannot run is refused by the scan: skip that candidate
      if (o.pipeline == 1) {
        const int rc = sk_hip_scan(pots, npots, probe, npos, pos.data(), dets, ndets, outs.ptrs.data(), 1, &o);
        if (rc == SK_ERR_INVALID) continue;
        if (rc != SK_OK) fail(rc, std::string("autotune: ") + last_error());
      }
      secs = scan_seconds(pots, npots, probe, npos, pos.data(), dets, ndets, outs.ptrs.data(), o, reps);
      const int k = report->ncandidates++;
      report->candidates[k] = o;
      report->rates[k] = (double)npos * (double)slices / secs;
      if (best < 0 || report->rates[k] > report->rates[best]) best = k;
    }
    if (best < 0) fail(SK_ERR_DEVICE, "autotune: no candidate ran");
    report->best = report->candidates[best];
    report->best_rate = report->rates[best];
    describe_device(report->device, sizeof(report->device));
  });
}

extern "C" int sk_hip_device_name(char* name, int len) {
  using namespace sk;
  return guarded_call([&] {
    if (!name || len < 1) fail(SK_ERR_INVALID, "device_name: no buffer");
    if (sk_hip_available() != 1) fail(SK_ERR_UNAVAILABLE, std::string("device_name: ") + sk_hip_last_error());
    describe_device(name, (size_t)len);
  });
}
