// Floating-point operations per pair of the influence and particle kernels,
// counted from the restated formulas (influence.h, stretch.h, vp.hip) as the
// kernels evaluate them.  Convention: every add, subtract, multiply and
// divide is one flop; each transcendental (sqrt, atan2, log, erf, exp) is
// also one flop here and is counted again separately in the *_trans
// constant, because on the device it is a software sequence of tens of
// instructions (FP64) and the flop rate alone understates the work.
// Compares and the far/near decision are not counted.  The throughput a
// kernel reaches is reported both as these flops per second against the
// VALU peak and, from the profiler's VALU instruction counters, as issued
// instructions against the issue peak.
#pragma once

namespace cpgpu {
namespace flops {

// ---- influence assembly (bodyPanel::panelPhiInf, bodyPanel.cpp:132-190)
// far pair, point source + point doublet (assemble far pass):
//   offset 3, r^2 5, sqrt, PN 5, source 2 (1 mul, 1 div), doublet 5
//   (r^3 2, 4pi 1, PN*area 1, div 1)
constexpr int kFarPair = 21;
constexpr int kFarPairTrans = 1;   // sqrt
constexpr int kFarPairDiv = 2;
// near pair, triangle (near_phi_f64): centre offset 3, r 5 + sqrt, PN 5;
//   three vertices: offset 3 + distance 5 + sqrt; three edges: edge 3,
//   s x a 9, Al 5, vortexPhi 42 + atan2 (a x s 9, l x (a x s) 9, PA 5, sm 5,
//   PB 2, num 5, den 7), doublet sum 1, |s| 5 + sqrt, A+B-S 2, GL 5 + log
//   (1/S, ratio, log, product), source sum 4; final scaling 2
constexpr int kNearTriPair = 280;
constexpr int kNearTriPairTrans = 13;  // 7 sqrt, 3 atan2, 3 log
constexpr int kNearTriPairDiv = 8;
// wake pair, quadrilateral doublet (panel::dubPhiInf, panel.cpp:208-252):
// far 18 + sqrt; near 286 + 9 (sqrt 5, atan2 4)
constexpr int kWakeFarPair = 19;
constexpr int kWakeNearQuadPair = 295;

// ---- panel -> point velocity (bodyPanel::panelV, bodyPanel.cpp:122-130,
// 305-348; panel::dubVInf, panel.cpp:666-711), source and doublet of one
// body panel times their strengths
//   far: offset 3, r^2 5, sqrt, r^3 and 4pi 3, area/den 1, source 3, point
//   doublet 23, strengths 12
constexpr int kP2PFarPair = 51;
constexpr int kP2PFarPairTrans = 1;
//   near triangle: per edge a, b, s 9, s x a and Al 14, vortexV 39 + 3 sqrt,
//   doublet sum 3, |a|, |b|, |s| 15 + 3 sqrt, GL 8 + log, vortexPhi 42 +
//   atan2, sm, sl 10, source sum 18; plus offset, PN, scaling 19, strengths 12
constexpr int kP2PNearTriPair = 529;
constexpr int kP2PNearTriPairTrans = 24;

// ---- particle kernels (particle.cpp:31-140), one (target, source) pair
// algebraic velocity (velInflAlgSmooth): offset 3, r 5 + sqrt, core terms 10
// + sqrt, cross product and accumulation 15
constexpr int kAlgPair = 35;
constexpr int kAlgPairTrans = 2;
// Gaussian velocity (velInfl): symmetric core 6 + sqrt, offset 3, |x| 5 +
// sqrt, rho 1, K 12 + erf + exp, 1/sigma^3 4, cross product 15
constexpr int kGaussPair = 50;
constexpr int kGaussPairTrans = 4;
// vortex stretching (vortexStretching): offset 3, |x| 5 + sqrt, core 6 +
// sqrt, xi 3 + exp, G 4 + erf, K 3, F 4, c 3, d.alpha 5, M alpha 12,
// 1/sigma^3 3, alpha_t x (M alpha) 15
constexpr int kStretchPair = 71;
constexpr int kStretchPairTrans = 4;
// viscous diffusion (viscousDiffusion): offset 3, core 6 + sqrt, rho 6 +
// sqrt, volumes 6, xi 7 + exp, factor 4, accumulation 15
constexpr int kDiffPair = 50;
constexpr int kDiffPairTrans = 3;

// ---- the configured case's forms (vp.hip), FP32 or FP64 arithmetic
// algebraic fixed-core velocity (kc_vel): offset 3, r^2 5, q 1, r^2 + 2.5
// s^2 2, q^2 sqrt(q) 2 + sqrt, divide 1, constant 1, cross product 9,
// scale 3, accumulation 3
constexpr int kAlgCorePair = 31;
constexpr int kAlgCorePairTrans = 1;
// its stretching (kc_stretch): offset 3, r^2 5, q 1, sqrt, f 6, h 8,
// (alpha_t . x) h 6, two cross products with their scales and the sum and
// accumulation 30
constexpr int kStretchCorePair = 60;
constexpr int kStretchCorePairTrans = 1;
// far panel -> point velocity in FP32 (p2p_far_f32): offset 3, r^2 5,
// rsqrt, 1/r^2 1, area / r^3 2, 3 (x . n) / r^2 7, doublet 13, source 8
constexpr int kP2PFarPairF32 = 40;
constexpr int kP2PFarPairF32Trans = 1;

}  // namespace flops
}  // namespace cpgpu
