// cf_tb_ring.hpp -- the temporally blocked RK4 step of one work unit with the state
// planes streamed into an LDS ring by direct global-to-LDS loads.
//
// The step body of tb_unit (cf_tb_body.hpp) keeps a register history of the state
// planes and loads plane P+2 into registers two iterations ahead. Here the state
// planes go from HBM straight into a ring of R = D + 2 LDS slots (one slot: phi
// and pi of the 32 x 32 region), D planes ahead of use, without passing through
// registers: each wavefront issues one 16-byte-per-lane global_load_lds per plane
// (wavefront w: rows 4(w&7) .. 4(w&7)+3 of component w>>3, two columns per lane).
// Iteration j (plane P = a0 - 4 + j) then reads x(P) from its slot into the
// register history and takes the b+-1 neighbours of stage 1's input x(P-1)
// straight from that plane's slot (no copy into a stage buffer). Everything else
// -- the stage schedule, the register histories of the stage inputs and partial
// sums, the stage buffers of y1..y3, the deferred buffer-resource stores and the
// per-site arithmetic -- is that of tb_unit, so the results are bit-identical.
//
// Ordering. The loads write LDS asynchronously and are counted by vmcnt like every
// vector memory operation. Each iteration issues, in this order: the two stores
// of the state finished two iterations earlier, then (after the barrier) the load
// of plane P + D. A prologue issues D groups of [two dropped stores, one load], so
// at the wait of iteration j exactly TBR_KV = 3 (D - 1) + 2 vector memory operations
// were issued after the load of plane P: s_waitcnt vmcnt(TBR_KV) completes it. The
// barrier that follows makes every wavefront's part of the plane visible (and the
// stage buffers, written before it). A slot is overwritten by the load issued
// after the barrier of iteration j + 2, by which time every wavefront has finished
// iteration j + 1, the last to read it. The barrier is a bare s_barrier with
// explicit waits: a workgroup fence would wait for every load in flight
// (vmcnt(0)) and undo the prefetch. Compiler barriers (empty asm with a memory
// clobber) keep the memory operations of a group in program order.
#pragma once

#include "cf_tb_body.hpp"

namespace cf {

typedef __attribute__((address_space(3))) void tb_lds_void;

// s_waitcnt immediate (gfx9 encoding): vmcnt(v) lgkmcnt(0), expcnt not waited on
constexpr int tb_waitcnt_vm_lgkm0(int v) { return (v & 15) | ((v >> 4) << 14) | (7 << 4); }

template <int D>
struct TbRing {
  static_assert(D >= 1 && D <= 16, "prefetch depth");
  static constexpr int R = D + 2;                     // slots
  static constexpr int PLANE = 2 * TBR * TBR;         // doubles per slot (phi, pi)
  static constexpr int X_BYTES = R * PLANE * 8;
  static constexpr int Y_BYTES = 2 * 3 * TBR * TBR * 8;  // stage inputs y1..y3, two sets
  static constexpr int BYTES = X_BYTES + Y_BYTES;
  static constexpr int KV = 3 * (D - 1) + 2;           // vm operations after the awaited load
};

// FP64 Klein-Gordon (+ lambda phi^3) step of one unit; smem holds TbRing<D>::BYTES.
template <bool LAM, int D>
__device__ __forceinline__ void tb_unit_ring(char *smem, const TbUnit &u, const TbStep &k,
                                             const double *__restrict__ xin, double *__restrict__ xout) {
  typedef TbRing<D> G;
  double *xr = reinterpret_cast<double *>(smem);               // [R][2][TBR][TBR]
  double *yb = reinterpret_cast<double *>(smem + G::X_BYTES);  // [2][3][TBR][TBR]
  const int tid = threadIdx.x;
  const int r = tid >> 5, cc = tid & 31;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6), ln = tid & 63;  // wavefront index (uniform)
  const int64_t N = u.N, NN = N * N, NA = u.NA, N3 = NN * NA;  // N3: offset of the pi planes
  const int plane_bytes = (int)(NN * (int64_t)sizeof(double));
  int64_t b = (u.b0 - TBH + r) % N;
  if (b < 0) b += N;
  int64_t c = (u.c0 - TBH + cc) % N;
  if (c < 0) c += N;
  const int64_t col = b * N + c;
  const bool out_ok = r >= TBH && r < TBH + u.T && u.b0 + (r - TBH) < N && cc >= TBH && cc < TBH + u.Tc &&
                      u.c0 + (cc - TBH) < N;
  const int rm = r > 0 ? r - 1 : 0, rp = r < TBR - 1 ? r + 1 : TBR - 1;

  // this lane's 16 bytes of a plane load: component w >> 3, region row 4(w&7) + (ln>>4),
  // region columns 2(ln&15), +1 (an even column and its successor: N is even, so the
  // pair never straddles the periodic wrap)
  const int comp = w >> 3;
  const int lrow = 4 * (w & 7) + (ln >> 4), lcol = 2 * (ln & 15);
  int64_t lb = (u.b0 - TBH + lrow) % N;
  if (lb < 0) lb += N;
  int64_t lc = (u.c0 - TBH + lcol) % N;
  if (lc < 0) lc += N;
  const int64_t lane_el = lb * N + lc;  // the lane's first element within a plane
  const int lds_wave = comp * TBR * TBR + 4 * (w & 7) * TBR;  // first element this wavefront writes in a slot

  auto wrap_plane = [&](int64_t P) -> int64_t {
    int64_t q = P % NA;
    if (q < 0) q += NA;
    return q * NN;
  };
  int64_t pl_ld = wrap_plane(u.a0 - TBH);  // next plane to load
  int s_ld = 0;                            // its slot
  auto issue = [&]() {
    // (uniform plane offset + per-lane offset: a precomputed per-lane pointer costs registers)
    __builtin_amdgcn_global_load_lds((const void *)(xin + (int64_t)comp * N3 + pl_ld + lane_el),
                                     (tb_lds_void *)(xr + s_ld * G::PLANE + lds_wave), 16, 0, 0);
    pl_ld += NN;
    if (pl_ld == N3) pl_ld = 0;
    s_ld = (s_ld + 1 == G::R) ? 0 : s_ld + 1;
  };

  constexpr int TBU = 8;
  double XF[TBU] = {}, XP[TBU] = {};
  double Y1F[TBU] = {}, Y1P[TBU] = {}, Y2F[TBU] = {}, Y2P[TBU] = {}, Y3F[TBU] = {}, Y3P[TBU] = {};
  double AF[TBU] = {}, AP[TBU] = {};
  const int iters = (int)(u.a1 - u.a0) + 8;
  TbPending<double> Q[2] = {};
  Q[0].off = Q[1].off = TB_OOB;
  auto flush = [&](const TbPending<double> &q) {
    tb_store(q.f, tb_plane(xout + q.plane, plane_bytes), q.off);
    tb_store(q.p, tb_plane(xout + N3 + q.plane, plane_bytes), q.off);
  };
  for (int g = 0; g < D; ++g) {  // prologue: D groups [two dropped stores, one load]
    asm volatile("" ::: "memory");
    flush(Q[0]);
    issue();
  }
  int64_t off_st = wrap_plane(u.a0 - TBH - 4);  // plane P-4
  int s0 = 0, s1 = G::R - 1;                    // slots of P and P-1

  auto iteration = [&](auto slot, const int j) {
    constexpr int S = decltype(slot)::value;  // j & 1: stage-buffer set, pending store
    constexpr int I0 = 0, I1 = TBU - 1, I2 = TBU - 2, I3 = TBU - 3;
    double(*buf)[TBR][TBR] = reinterpret_cast<double(*)[TBR][TBR]>(yb + S * 3 * TBR * TBR);
    TbPending<double> &q = Q[S];
    flush(q);  // the new state of iteration j-2
    const uint32_t st_off = ((j >= 8) && (j < iters) && out_ok) ? (uint32_t)(col * (int64_t)sizeof(double)) : TB_OOB;
    buf[0][r][cc] = Y1F[I1];
    buf[1][r][cc] = Y2F[I1];
    buf[2][r][cc] = Y3F[I1];
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(tb_waitcnt_vm_lgkm0(G::KV));
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    issue();  // plane P + D
    const double(*x0)[TBR] = reinterpret_cast<const double(*)[TBR]>(xr + s0 * G::PLANE);
    const double(*x1)[TBR] = reinterpret_cast<const double(*)[TBR]>(xr + s1 * G::PLANE);
    XF[I0] = x0[r][cc];
    XP[I0] = x0[TBR + r][cc];

    // ---- stage 1 at P-1: k1 = f(x)
    double kf = XP[I1];
    double kp = tb_rhs(XF[I1], XF[I0], XF[I2], x1[rp][cc], x1[rm][cc], tb_lane_next(XF[I1]), tb_lane_prev(XF[I1]),
                       XP[I1], k.c_h[0], k.c_m, k.lam, k.inv[0], LAM);
    Y1F[I0] = XF[I1] + k.dh * kf;
    Y1P[I0] = XP[I1] + k.dh * kp;
    AF[I0] = XF[I1] + k.dt6 * kf;
    AP[I0] = XP[I1] + k.dt6 * kp;
    // ---- stage 2 at P-2: k2 = f(y1)
    kf = Y1P[I1];
    kp = tb_rhs(Y1F[I1], Y1F[I0], Y1F[I2], buf[0][rp][cc], buf[0][rm][cc], tb_lane_next(Y1F[I1]),
                tb_lane_prev(Y1F[I1]), Y1P[I1], k.c_h[1], k.c_m, k.lam, k.inv[1], LAM);
    Y2F[I0] = XF[I2] + k.dh * kf;
    Y2P[I0] = XP[I2] + k.dh * kp;
    AF[I1] = AF[I1] + k.dt3 * kf;
    AP[I1] = AP[I1] + k.dt3 * kp;
    // ---- stage 3 at P-3: k3 = f(y2)
    kf = Y2P[I1];
    kp = tb_rhs(Y2F[I1], Y2F[I0], Y2F[I2], buf[1][rp][cc], buf[1][rm][cc], tb_lane_next(Y2F[I1]),
                tb_lane_prev(Y2F[I1]), Y2P[I1], k.c_h[2], k.c_m, k.lam, k.inv[2], LAM);
    Y3F[I0] = XF[I3] + k.dt * kf;
    Y3P[I0] = XP[I3] + k.dt * kp;
    AF[I2] = AF[I2] + k.dt3 * kf;
    AP[I2] = AP[I2] + k.dt3 * kp;
    // ---- stage 4 at P-4: k4 = f(y3); new state
    kf = Y3P[I1];
    kp = tb_rhs(Y3F[I1], Y3F[I0], Y3F[I2], buf[2][rp][cc], buf[2][rm][cc], tb_lane_next(Y3F[I1]),
                tb_lane_prev(Y3F[I1]), Y3P[I1], k.c_h[3], k.c_m, k.lam, k.inv[3], LAM);
    q.f = AF[I3] + k.dt6 * kf;
    q.p = AP[I3] + k.dt6 * kp;
    q.off = st_off;
    q.plane = off_st;
    off_st += NN;
    if (off_st == N3) off_st = 0;
    s1 = s0;
    s0 = (s0 + 1 == G::R) ? 0 : s0 + 1;
  };
  // advance one plane: the value d planes behind P moves from index -d to -(d+1) (mod TBU)
  auto rotate = [](double(&a)[TBU]) {
    const double t = a[0];
#pragma unroll
    for (int i = 0; i + 1 < TBU; ++i) a[i] = a[i + 1];
    a[TBU - 1] = t;
  };
  auto advance = [&]() {
    rotate(XF);
    rotate(XP);
    rotate(Y1F);
    rotate(Y1P);
    rotate(Y2F);
    rotate(Y2P);
    rotate(Y3F);
    rotate(Y3P);
    rotate(AF);
    rotate(AP);
  };
  using std::integral_constant;
  int j = 0;
  for (; j + 1 < iters; j += 2) {
    iteration(integral_constant<int, 0>{}, j);
    advance();
    iteration(integral_constant<int, 1>{}, j + 1);
    advance();
  }
  if (j < iters) iteration(integral_constant<int, 0>{}, j);
  asm volatile("" ::: "memory");
  flush(Q[0]);
  flush(Q[1]);
  __builtin_amdgcn_s_waitcnt(tb_waitcnt_vm_lgkm0(0));  // no load into LDS outlives the unit
}

}  // namespace cf
