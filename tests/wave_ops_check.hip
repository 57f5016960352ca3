// GPU check of the wavefront helpers in ugu_amd/csrc/device_common.h that the
// kernels build their scans and reductions on (DPP row shifts and row
// broadcasts): 32 wavefronts of seeded data, every lane's result compared on
// the host with a sequential computation. Built by ugu_amd/csrc/Makefile,
// run by tests/test_gpu_wave_ops.py.
//
// usage: wave_ops_check; prints "wave ops: <n> lanes checked, <m> mismatches"
// and exits 1 on any mismatch.

#include <cstdint>
#include <cstdio>
#include <vector>

#include "../ugu_amd/csrc/device_common.h"

namespace {

constexpr int kWaves = 32;
constexpr int kOut = 7;  // results per lane

__global__ void k_wave_ops(const int* in, const uint32_t* inu, const int* runs, int* out,
                           float* fout) {
  // one wavefront per 64-thread block, all lanes active
  const int i = blockIdx.x * 64 + threadIdx.x;
  int* o = out + static_cast<int64_t>(i) * kOut;
  o[0] = ugu_hip::wave_incl_sum(in[i]);
  o[1] = ugu_hip::wave_incl_max(in[i]);
  o[2] = static_cast<int>(ugu_hip::wave_max_u32(inu[i]));
  o[3] = static_cast<int>(ugu_hip::wave_min_u32(inu[i]));
  o[4] = ugu_hip::wave_prev(in[i], -7);
  o[5] = ugu_hip::wave_next(in[i], -9);
  o[6] = 0;
  fout[i] = ugu_hip::run_sum(static_cast<float>(in[i]), runs[i]);
}

uint32_t lcg(uint32_t& s) {
  s = s * 1664525u + 1013904223u;
  return s;
}

}  // namespace

int main() {
  const int n = kWaves * 64;
  std::vector<int> in(n), runs(n), out(size_t(n) * kOut);
  std::vector<uint32_t> inu(n);
  std::vector<float> fout(n);
  uint32_t s = 2024u;
  for (int w = 0; w < kWaves; w++) {
    // wave 0: all zeros; wave 1: one run over the whole wave; others random
    for (int l = 0; l < 64; l++) {
      const int i = w * 64 + l;
      in[i] = w == 0 ? 0 : static_cast<int>((lcg(s) >> 8) % 1000);
      inu[i] = w == 0 ? 0u : (w == 2 && l == 37 ? 0xffffffffu : lcg(s));
    }
    int id = 0;
    for (int l = 0; l < 64;) {
      const uint32_t r = lcg(s);
      if (w == 1) {
        runs[w * 64 + l] = 5;
        l++;
        continue;
      }
      if ((r >> 5) % 5 == 0) {  // a lane of its own (as invalid items are)
        runs[w * 64 + l] = -1 - l;
        l++;
        continue;
      }
      const int len = 1 + static_cast<int>((r >> 10) % 24);
      for (int k = 0; k < len && l < 64; k++) runs[w * 64 + l++] = id;
      id++;
    }
  }
  int *d_in, *d_runs, *d_out;
  uint32_t* d_inu;
  float* d_fout;
  if (hipMalloc(&d_in, n * 4) || hipMalloc(&d_runs, n * 4) || hipMalloc(&d_inu, n * 4) ||
      hipMalloc(&d_out, size_t(n) * kOut * 4) || hipMalloc(&d_fout, n * 4)) {
    std::printf("hipMalloc failed\n");
    return 2;
  }
  if (hipMemcpy(d_in, in.data(), n * 4, hipMemcpyHostToDevice) ||
      hipMemcpy(d_runs, runs.data(), n * 4, hipMemcpyHostToDevice) ||
      hipMemcpy(d_inu, inu.data(), n * 4, hipMemcpyHostToDevice))
    return 2;
  hipLaunchKernelGGL(k_wave_ops, dim3(kWaves), dim3(64), 0, 0, d_in, d_inu, d_runs, d_out, d_fout);
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(out.data(), d_out, size_t(n) * kOut * 4, hipMemcpyDeviceToHost) ||
      hipMemcpy(fout.data(), d_fout, n * 4, hipMemcpyDeviceToHost)) {
    std::printf("kernel or copy failed\n");
    return 2;
  }
  long checked = 0, bad = 0;
  auto fail = [&](const char* what, int w, int l, long got, long want) {
    if (bad < 10) std::printf("%s wave %d lane %d: got %ld want %ld\n", what, w, l, got, want);
    bad++;
  };
  for (int w = 0; w < kWaves; w++) {
    const int b = w * 64;
    uint32_t mx = 0, mn = 0xffffffffu;
    for (int l = 0; l < 64; l++) {
      mx = inu[b + l] > mx ? inu[b + l] : mx;
      mn = inu[b + l] < mn ? inu[b + l] : mn;
    }
    int sum = 0, rmax = 0;
    for (int l = 0; l < 64; l++) {
      const int i = b + l;
      const int* o = &out[size_t(i) * kOut];
      sum += in[i];
      rmax = in[i] > rmax ? in[i] : rmax;
      if (o[0] != sum) fail("incl_sum", w, l, o[0], sum);
      if (o[1] != rmax) fail("incl_max", w, l, o[1], rmax);
      if (static_cast<uint32_t>(o[2]) != mx) fail("max_u32", w, l, static_cast<uint32_t>(o[2]), mx);
      if (static_cast<uint32_t>(o[3]) != mn) fail("min_u32", w, l, static_cast<uint32_t>(o[3]), mn);
      const int prev = l == 0 ? -7 : in[i - 1], next = l == 63 ? -9 : in[i + 1];
      if (o[4] != prev) fail("prev", w, l, o[4], prev);
      if (o[5] != next) fail("next", w, l, o[5], next);
      // segmented inclusive sum over the lane's run (small integers: exact)
      float want = 0.0f;
      for (int j = l; j >= 0 && runs[b + j] == runs[i]; j--) want += static_cast<float>(in[b + j]);
      if (fout[i] != want) fail("run_sum", w, l, static_cast<long>(fout[i]), static_cast<long>(want));
      checked++;
    }
  }
  std::printf("wave ops: %ld lanes checked, %ld mismatches\n", checked, bad);
  return bad == 0 ? 0 : 1;
}
