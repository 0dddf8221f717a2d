// Common device-side declarations: error checking, launch helpers and the
// per-geometry device tables.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "influence.h"

namespace cpgpu {

// Error carried out of a HIP call; converted to cp_status at the C ABI.
struct HipError {
    hipError_t code;
    std::string where;
};

#define CP_HIP(call)                                                             \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) throw ::cpgpu::HipError{e_, std::string(#call)};   \
    } while (0)

#define CP_LAUNCH_CHECK() CP_HIP(hipGetLastError())

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Device tables of one geometry (all owned by the geometry object).
struct DevGeom {
    int n = 0;        // body panels
    int nw = 0;       // wake panels (first row)
    int64_t lda = 0;  // leading dimension of A, B (rows padded to 256)
    PanelRec* body = nullptr;     // n records
    PanelRec* wake = nullptr;     // nw records
    double* cx = nullptr;         // collocation points (= body centres), lda each
    double* cy = nullptr;
    double* cz = nullptr;
    // Kutta folding by body column: for column c, entries [ptr[c], ptr[c+1])
    int* fold_ptr = nullptr;
    int* fold_w = nullptr;
    double* fold_c = nullptr;
    // bounding spheres of the 64-row tiles of collocation points (centre,
    // radius: 4 per tile), for the near-field pass of the assembly
    double* tsph = nullptr;
    int ntiles = 0;
};

// ---- kernel launchers (assemble.hip) ----
void launch_wake_influence(const DevGeom& g, double* W, int64_t ldw, hipStream_t s);
void launch_assemble_f64(const DevGeom& g, const double* W, int64_t ldw, double* A, double* B, hipStream_t s);
void launch_assemble_f32(const DevGeom& g, const double* W, int64_t ldw, float* A, float* B, hipStream_t s);
// FP64 near-field corrections dA = A64 - A32, dB = B64 - B32 on the listed
// (row, column) entries of the FP32 matrices.
void launch_near_correction(const DevGeom& g, const double* W, int64_t ldw, const float* A32, const float* B32,
                            int nnz, const int* rows, const int* cols, double* dA, double* dB, hipStream_t s);

// ---- near-field pairs and tiles on the device (nearcsr.hip) ----
// Pairs |c_i - c_j| <= factor * longSide_j of the body panels: reach[t] =
// factor * the largest longSide of 64-panel tile t (g.ntiles doubles); then
// the counts per column chunk (cnt: near_chunks(g) * n ints, left holding
// each chunk's offset within its row) and per row (total: n ints); then,
// from rowptr = exclusive scan of total, each row's columns ascending in
// cols and its row index in rows (rowptr[n] entries).
void launch_near_reach(const DevGeom& g, double factor, double* reach, hipStream_t s);
int near_chunks(const DevGeom& g);
void launch_near_count(const DevGeom& g, double factor, const double* reach, int* cnt, int* total, hipStream_t s);
void launch_near_fill(const DevGeom& g, double factor, const double* reach, const int* cnt, const int* rowptr,
                      int* cols, int* rows, hipStream_t s);
// out[0] = 0, out[t + 1] = in[0] + ... + in[t] (n + 1 ints; one workgroup)
void launch_exclusive_scan(const int* in, int n, int* out, hipStream_t s);
// The 32-row x 16-column tiles holding a pair: mark (nrt x nkb bytes,
// zeroed by the caller) and cnt[rt] = row tile rt's count (nrt ints), then
// from tptr = exclusive scan of cnt, each row tile's column blocks ascending.
void launch_near_tile_count(int nnz, const int* rows, const int* cols, int nrt, int nkb, unsigned char* mark,
                            int* cnt, hipStream_t s);
void launch_near_tile_fill(const unsigned char* mark, int nrt, int nkb, const int* tptr, int* tkb, hipStream_t s);

}  // namespace cpgpu
