// FP64 single right-hand-side Krylov solver and GEMV (krylov.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace cpgpu {

// slots of the device scalar block
enum { SC_RR = 0, SC_R0R = 1, SC_ALPHA = 2, SC_OMEGA = 3 };

struct SolveStats {
    int iters = 0;
    int restarts = 0;
    int refinements = 0;
    double rel_residual = 0;
};

// y = sign * A x for a column-major FP64 matrix (split-K GEMV).
class Gemv64 {
public:
    Gemv64(int n, int64_t lda);
    int slices() const { return S_; }
    void partial(const double* A, const double* x, double* part, hipStream_t s) const;
    void apply(const double* A, const double* x, double* y, double sign, double* part, hipStream_t s) const;

private:
    int n_;
    int64_t lda_;
    int S_ = 1, chunk_ = 1;
};

// Eigen-equivalent BiCGSTAB on the device (one RHS).
class Bicgstab64 {
public:
    Bicgstab64(int n, int64_t lda);
    ~Bicgstab64();
    Bicgstab64(const Bicgstab64&) = delete;
    Bicgstab64& operator=(const Bicgstab64&) = delete;
    // b, x: device vectors of length lda (x padded rows are zeroed).
    // invd: optional precomputed Jacobi diagonal (device), else from A.
    // x0: optional initial guess (device, length lda); null = zero, Eigen's
    // solve(); otherwise solveWithGuess (tolerance still relative to |b|)
    SolveStats solve(const double* A, const double* b, double* x, double tol, int maxit, hipStream_t s,
                     const double* invd = nullptr, const double* x0 = nullptr);
    double* gemv_scratch() { return part_; }
    const Gemv64& gemv() const { return gemv_; }

private:
    void fetch(hipStream_t s);
    static constexpr int kNumVec = 9;
    int n_;
    int64_t lda_;
    int nb_;
    Gemv64 gemv_;
    double* ws_ = nullptr;
    double* vec_[kNumVec] = {};
    double* part_ = nullptr;
    double* dp_[8] = {};
    double* sc_ = nullptr;
    double* host_sc_ = nullptr;
};

}  // namespace cpgpu
