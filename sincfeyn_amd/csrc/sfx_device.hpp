// sfx_device.hpp -- the library's device state and the graph-independent device operations, shared by
// the runtime translation units (sfx_runtime.hip, sfx_hsum.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include "sfx_integrand.hpp"
#include "../../include/sfx.h"

extern "C" const sfx_graph_entry *const SFX_GRAPHS[];
extern "C" const int SFX_NGRAPHS;

namespace sfx {

// error reporting: per calling thread, returned by sfx_last_error()
void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);
#define SFX_HIP(x) do{ hipError_t e_ = (x); if(e_ != hipSuccess) return sfx::hip_fail(e_, #x, __FILE__, __LINE__); }while(0)

struct Lib {
    bool init = false;
    int device = -1;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop;
    std::string info;
};
Lib &lib();
int require_init();

// a device buffer that grows on demand.  Bound to a stream (the context's), its memory is allocated and
// freed in that stream's order (hipMallocAsync / hipFreeAsync: no device-wide synchronisation, so a
// worker that grows a buffer does not wait for every other worker's kernels); unbound, with hipMalloc /
// hipFree.  reserve() does not keep the contents.
struct DevBuf {
    void *p = nullptr; size_t cap = 0; hipStream_t s = nullptr; bool bound = false;
    int reserve(size_t bytes);
    void release();
    template<class T> T *as() const { return (T*)p; }
};

// double-double accumulator as two doubles (device and host)
struct ddacc { double hi, lo; };

// per-tile partial of a refinement block (see k_level_partials)
struct LevelPart { double s[2], a[2], b1[2], b2[2], drop[2], u[2]; uint32_t nv, nskip, nkeep, nq, nd, nmiss; };
// level accumulator in device memory, added to by every block of the level in block order
struct LevelAcc {
    double s[2], a[2], b1[2], b2[2], drop[2], u[2];
    unsigned long long nv, nskip, nkept, nq, nd, nmiss;
};

constexpr int TILE_THREADS = 256;
constexpr int TILE_ITEMS = 32;                      // items per thread
constexpr long TILE = (long)TILE_THREADS*TILE_ITEMS;

// exclusive scan of the block counts cnt[0 .. nb) into off[], their total into *total (one workgroup)
hipError_t scan_blocks(hipStream_t s, const uint32_t *cnt, long nb, uint32_t *off, uint32_t *total);
// ordered compaction: out[0..*count-1] = the indices i < n with (fl[i] & mask) != 0, ascending.
// Scratch: blk (2 * ceil(n/TILE) + 1 uint32).
hipError_t compact_flags(hipStream_t s, const uint8_t *fl, long n, uint8_t mask, uint32_t *blk, uint32_t *out, uint32_t *count);

// the sums and the kept children of one refinement block (reference sum.c:726-746): for every valid
// child, in child order, S += v pw, A += |v|, skipped += skip; at the last level the shadow-tier sums;
// at other levels the children with |v| >= thr become cells of the next level (flags from the parent's
// skip bit, the child's near flag and tier, sum.c:744) and the others add |v| pw to the dropped weight.
// The double-double fallback results (e2, nd of the ddcnt fallback points) are reduced too.  Everything
// is added to *acc in a fixed order; nxt receives the kept cells at offset acc->nkept (before the call).
// e2[i] is read for the children whose flag has SFX_FL_DD (written there by the fallback kernel).
struct LevelArgs {
    const double *v; const uint8_t *fl; const sfx_cell_t *cells; long n; int D; int st; int last; int abskip;
    double pw, thr;
    const double *e2;
    sfx_cell_t *nxt; long nxt_cap;
    LevelPart *part; unsigned long long *keepoff;   // scratch: ceil(n/TILE) entries each
    LevelAcc *acc;
};
hipError_t level_reduce(hipStream_t s, const LevelArgs &a);

// the sums and the kept nodes of one block of a full-box scan (reference sum.c:166-191): over the n
// nodes of the block, S += v pw, A += |v| pw, the outer shells' |v| pw (flag bytes SFX_FL_SHELL1/2, in
// LevelAcc::b1/b2), and with keep, the dropped weight (|v| < thr) and the count of kept nodes (|v| >= thr);
// the fallback points' e2 and unresolved marks as in level_reduce.  Added to *acc in a fixed order;
// keepoff receives each tile's first output slot (from acc->nkept before the call).
struct ScanArgs {
    const double *v; const uint8_t *fl; const double *e2; const sfx_cell_t *nodes; long n;
    double pw, thr; int keep;
    sfx_cell_t *out; long out_cap;                  // the kept nodes' list (scan_keep)
    LevelPart *part; unsigned long long *keepoff;   // scratch: ceil(n/TILE) entries each
    LevelAcc *acc;
};
hipError_t scan_sums(hipStream_t s, const ScanArgs &a);
// the kept nodes of the block (flag byte 0, as the reference pushes them) to out[keepoff ...], in node
// order within the block; slots at or past out_cap are not written
hipError_t scan_keep(hipStream_t s, const ScanArgs &a);

}  // namespace sfx
