// Host-side internals shared by the translation units of libtsph: status
// handling, the growable device scratch arena and the context object.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/tsph.h"

struct tsph_ctx;

namespace tsph {

void set_error(const char* fmt, ...);

#define TSPH_CHECK_HIP(expr)                                                        \
  do {                                                                              \
    hipError_t err_ = (expr);                                                       \
    if (err_ != hipSuccess) {                                                       \
      ::tsph::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(err_),    \
                        __FILE__, __LINE__);                                        \
      return TSPH_ERR_HIP;                                                          \
    }                                                                               \
  } while (0)

#define TSPH_CHECK_LAUNCH() TSPH_CHECK_HIP(hipGetLastError())

#define TSPH_REQUIRE(cond, code, ...)                                               \
  do {                                                                              \
    if (!(cond)) {                                                                  \
      ::tsph::set_error(__VA_ARGS__);                                               \
      return code;                                                                  \
    }                                                                               \
  } while (0)

#define TSPH_TRY(expr)                                                              \
  do {                                                                              \
    int st_ = (expr);                                                               \
    if (st_ != TSPH_OK) return st_;                                                 \
  } while (0)

// A device buffer that only grows.  Growth happens outside the timed step
// loop (first call, or a call with more particles than before).
struct DevBuf {
  void* ptr = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap && ptr) return TSPH_OK;
    if (ptr) {
      hipError_t e = hipFree(ptr);
      if (e != hipSuccess) { set_error("hipFree failed: %s", hipGetErrorString(e)); return TSPH_ERR_HIP; }
    }
    ptr = nullptr; cap = 0;
    size_t want = bytes < 256 ? 256 : bytes;
    want += want / 8;   // headroom so slightly larger calls reuse the block
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
      set_error("hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
      ptr = nullptr;
      return TSPH_ERR_OOM;
    }
    cap = want;
    return TSPH_OK;
  }
  template <typename T> T* as() const { return static_cast<T*>(ptr); }
  void release() { if (ptr) (void)hipFree(ptr); ptr = nullptr; cap = 0; }
};

// Result of a linked-cell search: particles grouped by cell.
struct CellSearch {
  int64_t n = 0, ncells = 0;           // particles, occupied cells
  int dim = 0;
  int periodic = 0;
  int morton = 0;                      // keys are Morton codes of the cells (resident steppers)
  int64_t cmin[3] = {0, 0, 0};
  int64_t ext[3] = {1, 1, 1};
  int64_t stride[3] = {1, 1, 1};
  double width[3] = {0, 0, 0};         // periodic cell widths
  int64_t ncell_per[3] = {0, 0, 0};    // periodic cells per axis
  DevBuf cells;      // int64 (n, dim) cell coordinates (min-normalised)
  DevBuf keys;       // int64 (n) row-major key per particle
  DevBuf skeys;      // int64 (n) sorted keys
  DevBuf order;      // int64 (n) particle index per sorted slot (stable)
  DevBuf order32;    // int32 copy for the sort
  DevBuf ukeys;      // int64 (ncells) occupied keys ascending
  DevBuf ustart;     // int64 (ncells + 1) first sorted slot of each cell
  DevBuf flags;      // int32 scratch
  DevBuf minmax;     // uint64 2*3 encoded min/max
  DevBuf mm_blk;     // uint64 6 per block of the cell-coordinate pass
  DevBuf sort_tmp;   // rocprim temporary storage
  DevBuf idx_in;     // int32 0..n-1
};

int exclusive_scan(tsph_ctx* ctx, const unsigned long long* cnt, int64_t n, int64_t* off,
                   hipStream_t stream, int64_t* total);
template <typename T, int D>
int run_search(tsph_ctx* ctx, const T* pos, int64_t n, double cutoff, const double* box,
               hipStream_t stream, bool morton = false);

}  // namespace tsph

// Opaque handle of the C ABI.
struct tsph_ctx {
  tsph::CellSearch search;
  // pair export
  tsph::DevBuf pair_count;   // int64 (n + 1)
  tsph::DevBuf pair_off;     // int64 (n + 1)
  int64_t npairs = 0;
  tsph::DevBuf scan_tmp;
  // neighbour lists (CSR over targets), built from pairs or from cells
  tsph::DevBuf nb_off;       // int64 (ntargets + 1)
  tsph::DevBuf nb_idx;       // int32 (entries)
  tsph::DevBuf nb_cursor;    // int64 (ntargets)
  int64_t nb_targets = 0, nb_entries = 0;
  tsph::DevBuf wnb_off;      // CSR of wall targets -> fluid neighbours
  tsph::DevBuf wnb_idx;
  tsph::DevBuf wnb_cursor;
  int64_t wnb_targets = 0, wnb_entries = 0;
  // operator scratch
  tsph::DevBuf s_rho_raw, s_rho, s_p, s_aux, s_aux2, s_aux3, s_gr, s_flags, s_pack;
  tsph::DevBuf s_four, s_lap1;   // fourth-order viscous operator: I4 / 120 and the first Laplacian
  tsph::DevBuf s_cl;             // free-surface pressure closure: two pressure and two flag buffers
  tsph::DevBuf s_red;        // reductions
  tsph::DevBuf s_dp;         // delta-plus step intermediates
  tsph::DevBuf g_part, g_flag, g_idx, g_tmp, g_out;   // gauges: partials, near flags / rows, select scratch
};
