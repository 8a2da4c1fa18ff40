// Device-side graph ingest: the reference's read_edge + create_csr (multistep/scc_main.cpp:94-191)
// on the GPU.  The edge-list text is uploaded once; lines are found and parsed in parallel, and
// the CSR and transposed CSR are built with two stable radix sorts, so every adjacency list
// keeps the input order exactly as create_csr's counting sort does (duplicates and self-loops
// kept).
//
// Text format (read_edge): "n m\n", then m lines "u v".  read_edge takes each source with
// getline(' ') + atoi and each destination with getline('\n') + atoi, so a line is a source
// token up to the first space and a destination token up to the newline; atoi skips leading
// white space, takes an optional sign and stops at the first non-digit.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <string>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "msc_internal.h"

namespace msc {
namespace {

constexpr int kTextChunk = 16384;  // bytes of text per workgroup
constexpr int kHalo = 256;         // bytes past the chunk a line starting in it may use
constexpr int kPerThread = kTextChunk / kBlock;

enum : uint32_t { kErrRange = 1u, kErrFormat = 2u, kErrShort = 4u };

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }
__device__ __forceinline__ int wave_of() { return (int)(threadIdx.x >> 6); }

// inclusive prefix sum over the wave
__device__ __forceinline__ uint32_t wave_scan(uint32_t x) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane_id() >= d) x += y;
  }
  return x;
}

// workgroup sum, valid in thread 0
__device__ __forceinline__ unsigned long long group_sum(unsigned long long x, unsigned long long* s_red) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_down(x, d, 64);
  if (lane_id() == 0) s_red[wave_of()] = x;
  __syncthreads();
  unsigned long long t = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBlock / 64; ++i) t += s_red[i];
  return t;
}

// Newlines per chunk.
__global__ void __launch_bounds__(kBlock) k_count_lines(const unsigned char* __restrict__ text, int64_t bytes,
                                                        uint32_t* __restrict__ counts) {
  __shared__ unsigned long long s_red[kBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kTextChunk + (int64_t)threadIdx.x * kPerThread;
  unsigned long long c = 0;
  if (base < bytes) {
    const uint4* p = reinterpret_cast<const uint4*>(text + base);  // chunks are 16-byte aligned
#pragma unroll
    for (int q = 0; q < kPerThread / 16; ++q) {
      uint4 w = p[q];
      uint32_t words[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          int64_t at = base + q * 16 + k * 4 + b;
          c += (at < bytes && ((words[k] >> (8 * b)) & 0xffu) == '\n') ? 1u : 0u;
        }
    }
  }
  c = group_sum(c, s_red);
  if (threadIdx.x == 0) counts[blockIdx.x] = (uint32_t)c;
}

__device__ __forceinline__ bool is_space(unsigned c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

// atoi on s[i..): leading white space, optional sign, digits.  Returns the value (as int64,
// so an out-of-range id is caught rather than wrapped) and leaves i after the last digit.
__device__ __forceinline__ int64_t parse_int(const unsigned char* s, int& i, int end) {
  while (i < end && is_space(s[i])) ++i;
  bool neg = false;
  if (i < end && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
  int64_t v = 0;
  int digits = 0;
  while (i < end && s[i] >= '0' && s[i] <= '9' && digits < 12) {
    v = v * 10 + (s[i++] - '0');
    ++digits;
  }
  return neg ? -v : v;
}

// Parses the line starting at s[i] (its end within s[..end)): source token up to the first
// space, destination token up to the newline.
__device__ void parse_line(const unsigned char* s, int i, int end, int32_t n, int64_t line, int32_t* src, int32_t* dst,
                           uint32_t* err) {
  int sp = i;
  while (sp < end && s[sp] != ' ' && s[sp] != '\n') ++sp;
  if (sp >= end || s[sp] != ' ') {  // no space before the newline, or the line does not fit the halo
    atomicOr(err, kErrFormat);
    return;
  }
  int j = i;
  int64_t u = parse_int(s, j, sp);
  int k = sp + 1;
  int nl = k;
  while (nl < end && s[nl] != '\n') ++nl;
  if (nl >= end && end == kTextChunk + kHalo) {  // longer than the halo: not a line of two ids
    atomicOr(err, kErrFormat);
    return;
  }
  int64_t v = parse_int(s, k, nl);  // the last line of the text may lack its newline
  if (u < 0 || u >= n || v < 0 || v >= n) atomicOr(err, kErrRange);
  src[line] = (int32_t)u;
  dst[line] = (int32_t)v;
}

// Every line that starts inside the chunk: line L starts after the L-th newline (line 0 at
// the start of the text).  first[c] = newlines before chunk c.
__global__ void __launch_bounds__(kBlock) k_parse_lines(const unsigned char* __restrict__ text, int64_t bytes,
                                                        const uint32_t* __restrict__ first, int32_t n, int64_t m,
                                                        int32_t* __restrict__ src, int32_t* __restrict__ dst,
                                                        uint32_t* err) {
  __shared__ unsigned char s[kTextChunk + kHalo];
  __shared__ uint32_t s_wsum[kBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kTextChunk;
  const int64_t have = bytes - base < (int64_t)(kTextChunk + kHalo) ? bytes - base : (int64_t)(kTextChunk + kHalo);
  for (int i = threadIdx.x; i < kTextChunk + kHalo; i += kBlock) s[i] = i < have ? text[base + i] : (unsigned char)0;
  __syncthreads();
  const int end = (int)have;
  // newlines of this thread's slice and their rank within the chunk
  const int lo = threadIdx.x * kPerThread;
  uint32_t mine = 0;
  for (int i = lo; i < lo + kPerThread; ++i) mine += (i < end && s[i] == '\n') ? 1u : 0u;
  const uint32_t incl = wave_scan(mine);
  if (lane_id() == 63) s_wsum[wave_of()] = incl;
  __syncthreads();
  uint32_t rank = incl - mine;
  for (int w = 0; w < wave_of(); ++w) rank += s_wsum[w];
  const int64_t first_line = (int64_t)first[blockIdx.x];
  if (blockIdx.x == 0 && threadIdx.x == 0 && m > 0) parse_line(s, 0, end, n, 0, src, dst, err);
  for (int i = lo; i < lo + kPerThread && i < end; ++i) {
    if (s[i] != '\n') continue;
    const int64_t line = first_line + rank + 1;  // the line starting after this newline
    ++rank;
    if (line >= m) continue;
    if (i + 1 >= end) {  // the text ends right after this newline: fewer than m lines
      atomicOr(err, kErrShort);
      continue;
    }
    parse_line(s, i + 1, end, n, line, src, dst, err);
  }
}

// off[v] = first position of key v in the sorted keys (off[n] = m): each run boundary writes
// the offsets of the vertices between the previous key and its own.
__global__ void k_run_offsets(const int32_t* __restrict__ keys, int64_t m, int32_t n, uint32_t* __restrict__ off) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= m; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t prev = i == 0 ? -1 : (int64_t)keys[i - 1];
    const int64_t cur = i == m ? (int64_t)n : (int64_t)keys[i];
    for (int64_t v = prev + 1; v <= cur; ++v) off[v] = (uint32_t)i;
  }
}

__global__ void k_fill_off(uint32_t* off, int32_t n) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= n; v += (int64_t)gridDim.x * blockDim.x)
    off[v] = 0u;
}

unsigned grid_blocks(int64_t items) {
  int64_t b = (items + kBlock - 1) / kBlock;
  return (unsigned)(b < 1 ? 1 : (b > 8192 ? 8192 : b));  // the kernels stride over the rest
}

int bits_for(int32_t n) {
  int b = 0;
  while (b < 31 && (1LL << b) < (int64_t)n) ++b;
  return b > 0 ? b : 1;
}

}  // namespace

#define ING_TRY(expr)                                              \
  do {                                                             \
    hipError_t _e = (expr);                                        \
    if (_e != hipSuccess) {                                        \
      (void)hipGetLastError();                                     \
      rc = _e == hipErrorOutOfMemory ? MSC_E_OOM : MSC_E_DEVICE;   \
      goto done;                                                   \
    }                                                              \
  } while (0)

int ingest_text(hipStream_t st, const char* text, int64_t bytes, msc_csr* out) {
  // header "n m\n" on the host (getline(' ') + atoi, getline + strtoul)
  int64_t hdr = 0;
  while (hdr < bytes && text[hdr] != '\n') ++hdr;
  if (hdr >= bytes && bytes > 0) hdr = bytes;
  std::string head(text, (size_t)hdr);
  const size_t sp = head.find(' ');
  if (sp == std::string::npos) return MSC_E_INVALID;
  const long long n64 = std::atoll(head.substr(0, sp).c_str());
  const unsigned long long m64 = std::strtoull(head.c_str() + sp + 1, nullptr, 10);
  if (n64 < 0 || n64 > INT32_MAX || m64 >= (1ULL << 32)) return MSC_E_CAPACITY;
  const int32_t n = (int32_t)n64;
  const int64_t m = (int64_t)m64;
  const int64_t body = bytes > hdr ? bytes - hdr - 1 : 0;
  if (m > 0 && (n == 0 || body <= 0)) return MSC_E_INVALID;

  int rc = MSC_OK;
  unsigned char* dtext = nullptr;
  uint32_t *counts = nullptr, *first = nullptr, *err = nullptr, *oo = nullptr, *io = nullptr;
  int32_t *src = nullptr, *dst = nullptr, *keys = nullptr, *oa = nullptr, *ia = nullptr;
  void* tmp = nullptr;
  uint32_t herr = 0;
  const int64_t chunks = (body + kTextChunk - 1) / kTextChunk;
  const size_t padded = (size_t)(chunks * kTextChunk + kHalo);
  const size_t ma = (size_t)(m > 0 ? m : 1);
  ING_TRY(hipMalloc(&oo, ((size_t)n + 1) * 4));
  ING_TRY(hipMalloc(&io, ((size_t)n + 1) * 4));
  ING_TRY(hipMalloc(&oa, ma * 4));
  ING_TRY(hipMalloc(&ia, ma * 4));
  if (m == 0) {
    hipLaunchKernelGGL(k_fill_off, dim3(grid_blocks(n + 1)), dim3(kBlock), 0, st, oo, n);
    hipLaunchKernelGGL(k_fill_off, dim3(grid_blocks(n + 1)), dim3(kBlock), 0, st, io, n);
    ING_TRY(hipGetLastError());
    ING_TRY(hipStreamSynchronize(st));
  } else {
    ING_TRY(hipMalloc(&dtext, padded));
    ING_TRY(hipMalloc(&counts, (size_t)chunks * 4));
    ING_TRY(hipMalloc(&first, (size_t)chunks * 4));
    ING_TRY(hipMalloc(&err, 4));
    ING_TRY(hipMalloc(&src, ma * 4));
    ING_TRY(hipMalloc(&dst, ma * 4));
    ING_TRY(hipMalloc(&keys, ma * 4));
    ING_TRY(hipMemsetAsync(dtext + body, 0, padded - (size_t)body, st));
    ING_TRY(hipMemcpyAsync(dtext, text + hdr + 1, (size_t)body, hipMemcpyHostToDevice, st));
    ING_TRY(hipMemsetAsync(err, 0, 4, st));
    hipLaunchKernelGGL(k_count_lines, dim3((unsigned)chunks), dim3(kBlock), 0, st, dtext, body, counts);
    ING_TRY(hipGetLastError());
    {
      size_t need = 0;
      ING_TRY(rocprim::exclusive_scan(nullptr, need, counts, first, 0u, (size_t)chunks, rocprim::plus<uint32_t>(),
                                      st));
      size_t need2 = 0, need3 = 0;
      const int bits = bits_for(n);
      ING_TRY(rocprim::radix_sort_pairs(nullptr, need2, src, keys, dst, oa, (size_t)m, 0, bits, st));
      ING_TRY(rocprim::radix_sort_pairs(nullptr, need3, dst, keys, src, ia, (size_t)m, 0, bits, st));
      if (need2 > need) need = need2;
      if (need3 > need) need = need3;
      ING_TRY(hipMalloc(&tmp, need > 0 ? need : 1));
      size_t have = need;
      ING_TRY(rocprim::exclusive_scan(tmp, have, counts, first, 0u, (size_t)chunks, rocprim::plus<uint32_t>(), st));
      hipLaunchKernelGGL(k_parse_lines, dim3((unsigned)chunks), dim3(kBlock), 0, st, dtext, body, first, n, m, src,
                         dst, err);
      ING_TRY(hipGetLastError());
      uint32_t tail[2] = {0u, 0u};  // newlines before the last chunk, newlines in it
      ING_TRY(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, st));
      ING_TRY(hipMemcpyAsync(&tail[0], first + chunks - 1, 4, hipMemcpyDeviceToHost, st));
      ING_TRY(hipMemcpyAsync(&tail[1], counts + chunks - 1, 4, hipMemcpyDeviceToHost, st));
      ING_TRY(hipStreamSynchronize(st));
      // lines in the text: one per newline, plus a last one without its newline
      const int64_t lines = (int64_t)tail[0] + tail[1] + (text[bytes - 1] != '\n' ? 1 : 0);
      if (herr || lines < m) {
        rc = MSC_E_INVALID;
        goto done;
      }
      // out-lists: sort (src -> dst) pairs by src, stably; in-lists: (dst -> src) by dst
      have = need;
      ING_TRY(rocprim::radix_sort_pairs(tmp, have, src, keys, dst, oa, (size_t)m, 0, bits, st));
      hipLaunchKernelGGL(k_run_offsets, dim3(grid_blocks(m + 1)), dim3(kBlock), 0, st, keys, m, n, oo);
      ING_TRY(hipGetLastError());
      have = need;
      ING_TRY(rocprim::radix_sort_pairs(tmp, have, dst, keys, src, ia, (size_t)m, 0, bits, st));
      hipLaunchKernelGGL(k_run_offsets, dim3(grid_blocks(m + 1)), dim3(kBlock), 0, st, keys, m, n, io);
      ING_TRY(hipGetLastError());
      ING_TRY(hipStreamSynchronize(st));
    }
  }
  *out = msc_csr{n, m, oo, oa, io, ia, MSC_DEVICE};
done:
  if (dtext) (void)hipFree(dtext);
  if (counts) (void)hipFree(counts);
  if (first) (void)hipFree(first);
  if (err) (void)hipFree(err);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  if (keys) (void)hipFree(keys);
  if (tmp) (void)hipFree(tmp);
  if (rc != MSC_OK) {
    if (oo) (void)hipFree(oo);
    if (io) (void)hipFree(io);
    if (oa) (void)hipFree(oa);
    if (ia) (void)hipFree(ia);
  }
  return rc;
}

}  // namespace msc
