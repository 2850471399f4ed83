// Vendor-library comparison for the sort rows of bench.py: rocPRIM's
// radix_sort_pairs on the same shapes (2^28 u32 keys + u32 index payload,
// 2^27 u64 keys of 42 bits + u32 payload), hipEvent median of 5 after one
// warm-up. Keys are uniform random (std::mt19937_64, seed 2); the output is
// checked sorted. Built by `python bench.py --build-vendor-sort` (the package
// build runs it too), run by `python bench.py` with the primitives; it prints
// one JSON object.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <random>
#include <rocprim/device/device_radix_sort.hpp>
#include <vector>

template <class K>
static bool run(size_t n, unsigned end_bit, const char* label, bool first) {
  std::vector<K> hk(n);
  std::mt19937_64 g(2);
  const K mask = end_bit >= sizeof(K) * 8 ? ~K(0) : K((K(1) << end_bit) - 1);
  for (auto& k : hk) k = K(g()) & mask;
  std::vector<unsigned> hv(n);
  for (size_t i = 0; i < n; ++i) hv[i] = unsigned(i);
  K *k_in = nullptr, *k_out = nullptr;
  unsigned *v_in = nullptr, *v_out = nullptr;
  void* d_tmp = nullptr;
  size_t tmp = 0;
  if (hipMalloc(&k_in, n * sizeof(K)) != hipSuccess || hipMalloc(&k_out, n * sizeof(K)) != hipSuccess ||
      hipMalloc(&v_in, n * 4) != hipSuccess || hipMalloc(&v_out, n * 4) != hipSuccess)
    return false;
  hipMemcpy(k_in, hk.data(), n * sizeof(K), hipMemcpyHostToDevice);
  hipMemcpy(v_in, hv.data(), n * 4, hipMemcpyHostToDevice);
  rocprim::radix_sort_pairs(nullptr, tmp, k_in, k_out, v_in, v_out, n, 0, end_bit);
  if (hipMalloc(&d_tmp, tmp) != hipSuccess) return false;
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  std::vector<float> ts;
  for (int r = 0; r < 6; ++r) {
    hipEventRecord(a);
    rocprim::radix_sort_pairs(d_tmp, tmp, k_in, k_out, v_in, v_out, n, 0, end_bit);
    hipEventRecord(b);
    hipEventSynchronize(b);
    float ms = 0;
    hipEventElapsedTime(&ms, a, b);
    if (r) ts.push_back(ms);
  }
  std::sort(ts.begin(), ts.end());
  std::vector<K> out(n);
  hipMemcpy(out.data(), k_out, n * sizeof(K), hipMemcpyDeviceToHost);
  const bool ok = std::is_sorted(out.begin(), out.end());
  const float ms = ts[ts.size() / 2];
  std::printf("%s\"%s\": {\"n\": %zu, \"key_bits\": %u, \"ms\": %.4f, \"keys_per_s\": %.4g, \"sorted\": %s}",
              first ? "" : ", ", label, n, end_bit, ms, n / (ms * 1e-3), ok ? "true" : "false");
  hipFree(k_in);
  hipFree(k_out);
  hipFree(v_in);
  hipFree(v_out);
  hipFree(d_tmp);
  return ok;
}

int main() {
  std::printf("{");
  bool ok = run<unsigned>(size_t(1) << 28, 32, "rocprim_sort_u32_pairs_2^28", true);
  ok = run<unsigned long long>(size_t(1) << 27, 42, "rocprim_sort_u64_morton42_pairs_2^27", false) && ok;
  std::printf("}\n");
  return ok ? 0 : 1;
}
