// Shared definitions of the evaluator: task lists and buffer tables.
//
// An evaluation is a sequence of launches; each launch covers one list of
// output tiles.  A tile is (task, row-tile); a task is one output block C
// (a leaf block of Y or a skeleton block of wskel/uskel) and carries the list
// of K-segments that accumulate into it:
//     C = beta*C + sum_seg op(A_seg) * B_seg
// where A_seg is a stored block (D, B or VT, optionally transposed) and B_seg
// a block of a column buffer (W, wskel or uskel).  Grouping every product by
// its destination block makes each launch free of write conflicts without
// atomics; the reference gets the same property from row-blocking
// (sympiler/DSInspector.cpp:179-181).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mx {

// stored-block buffers
enum MatBuf : int32_t { MAT_D = 0, MAT_B = 1, MAT_VT = 2 };
// column buffers (N x nrhs leaf-blocked, or skeleton rows x nrhs)
enum VecBuf : int32_t { VEC_W = 0, VEC_WSKEL = 1, VEC_USKEL = 2, VEC_Y = 3 };

struct alignas(16) Task {
    int64_t c_row;      // first row of the output block in its column buffer
    int32_t m;          // rows of the output block (its leading dimension)
    int32_t c_buf;      // VecBuf of C
    int32_t seg_begin;  // segments [seg_begin, seg_end)
    int32_t seg_end;
    int32_t beta;       // 1: accumulate into C, 0: overwrite
    int32_t flags;      // reserved
};

struct alignas(16) Seg {
    int64_t a_off;      // element offset of A in its stored-block buffer
    int64_t b_row;      // first row of the B block in its column buffer
    int32_t a_buf;      // MatBuf
    int32_t a_trans;    // 0: A is m x k (lda); 1: op(A) = A^T, A stored k x m (lda)
    int32_t lda;
    int32_t k;          // contraction length
    int32_t b_buf;      // VecBuf
    int32_t b_ld;       // rows of the B block (its leading dimension)
};

template <typename T>
struct Bufs {
    const T *mat[3];    // D, B, VT
    T *vec[4];          // W (read only), wskel, uskel, Y
};

// Work items of a launch are (tile, column-tile) pairs.  Workgroups are
// dealt round-robin to the 8 XCDs; remap the hardware id so each XCD gets a
// contiguous range of work items: the column tiles of one output tile then
// share the XCD's L2 for their A blocks.
__device__ __forceinline__ int xcd_remap(int bid, int total) {
    const int nx = 8;
    int q = total / nx, r = total % nx;
    int x = bid % nx, i = bid / nx;
    return x * q + (x < r ? x : r) + i;
}

}  // namespace mx
