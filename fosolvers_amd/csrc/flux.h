// Device-side physics shared by the residual, Jacobian and findEuler kernels: primitive
// state, boundary ghost states (src/foeuler/foeuler.f90:209-330) and the Roe flux
// (src/libfosolvers/FVM/Euler.f90:90-126).
#pragma once
#include <hip/hip_runtime.h>

namespace fos {

__device__ inline double dsqrt(double x) { return sqrt(x); }
__device__ inline float dsqrt(float x) { return sqrtf(x); }
__device__ inline double dpow(double x, double y) { return pow(x, y); }
__device__ inline float dpow(float x, float y) { return powf(x, y); }
__device__ inline double dabs(double x) { return fabs(x); }
__device__ inline float dabs(float x) { return fabsf(x); }

// Reciprocal and reciprocal square root from the hardware estimates (v_rcp_f64, v_rsq_f64) refined
// by two Newton-Raphson steps each: the error of the estimate is squared per step, so the result
// is within a few ulp of the correctly rounded value at a third of the instructions of the IEEE
// division and square root sequences.  FP32 uses the estimates directly (1 ulp).
__device__ inline double frcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ inline float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double frsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x * y, y, 1.0);  // y <- y + y (1 - x y^2) / 2
  y = fma(0.5 * y, e, y);
  e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
}
__device__ inline float frsqrt(float x) { return __builtin_amdgcn_rsqf(x); }

// Forward-mode dual number: value v and derivative d along one direction.  Pushed through the
// primitive, ghost-state and Roe-flux templates below it yields their exact directional
// derivative (the analytic Jacobian-vector product).  |x| takes the derivative sign(x) d with
// sign(0) = 0, the average of the two one-sided derivatives at the kink.
template <class T>
struct Dual {
  T v, d;
  Dual() = default;
  __host__ __device__ constexpr Dual(T a) : v(a), d(T(0)) {}
  __host__ __device__ constexpr Dual(T a, T b) : v(a), d(b) {}
};
template <class T> __device__ inline Dual<T> operator+(Dual<T> a, Dual<T> b) { return {a.v + b.v, a.d + b.d}; }
template <class T> __device__ inline Dual<T> operator-(Dual<T> a, Dual<T> b) { return {a.v - b.v, a.d - b.d}; }
template <class T> __device__ inline Dual<T> operator-(Dual<T> a) { return {-a.v, -a.d}; }
template <class T> __device__ inline Dual<T> operator*(Dual<T> a, Dual<T> b) { return {a.v * b.v, a.v * b.d + a.d * b.v}; }
template <class T> __device__ inline Dual<T> operator/(Dual<T> a, Dual<T> b) {
  const T q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
template <class T> __device__ inline Dual<T> operator+(Dual<T> a, T b) { return {a.v + b, a.d}; }
template <class T> __device__ inline Dual<T> operator+(T a, Dual<T> b) { return {a + b.v, b.d}; }
template <class T> __device__ inline Dual<T> operator-(Dual<T> a, T b) { return {a.v - b, a.d}; }
template <class T> __device__ inline Dual<T> operator-(T a, Dual<T> b) { return {a - b.v, -b.d}; }
template <class T> __device__ inline Dual<T> operator*(Dual<T> a, T b) { return {a.v * b, a.d * b}; }
template <class T> __device__ inline Dual<T> operator*(T a, Dual<T> b) { return {a * b.v, a * b.d}; }
template <class T> __device__ inline Dual<T> operator/(Dual<T> a, T b) { return {a.v / b, a.d / b}; }
template <class T> __device__ inline Dual<T> operator/(T a, Dual<T> b) { return Dual<T>(a) / b; }
template <class T> __device__ inline Dual<T> &operator*=(Dual<T> &a, Dual<T> b) { return a = a * b; }
template <class T> __device__ inline bool operator<(Dual<T> a, Dual<T> b) { return a.v < b.v; }
template <class T> __device__ inline bool operator>(Dual<T> a, Dual<T> b) { return a.v > b.v; }

template <class T> __device__ inline Dual<T> frcp(Dual<T> x) {
  const T r = frcp(x.v);
  return {r, -r * r * x.d};
}
template <class T> __device__ inline Dual<T> frsqrt(Dual<T> x) {
  const T y = frsqrt(x.v);
  return {y, T(-0.5) * y * y * y * x.d};
}
template <class T> __device__ inline Dual<T> dsqrt(Dual<T> x) {
  const T r = dsqrt(x.v);
  return {r, r > T(0) ? T(0.5) * x.d / r : T(0)};
}
template <class T> __device__ inline Dual<T> dpow(Dual<T> x, Dual<T> e) {
  const T p = dpow(x.v, e.v);
  T d = x.d != T(0) ? p * e.v * x.d / x.v : T(0);
  if (e.d != T(0)) d += p * e.d * (T)log((double)x.v);
  return {p, d};
}
template <class T> __device__ inline Dual<T> dabs(Dual<T> x) {
  return {x.v >= T(0) ? x.v : -x.v, x.v > T(0) ? x.d : (x.v < T(0) ? -x.d : T(0))};
}

// unit normal and area from the stored area-weighted normal (a = |a n|, n = (a n) / a; a zero area
// gives n = 0 and a = 0, i.e. a zero flux)
template <class TC>
__device__ inline void face_geom(TC ax, TC ay, TC az, TC &nx, TC &ny, TC &nz, TC &area) {
  const TC a2 = ax * ax + ay * ay + az * az;
  const TC yv = a2 > TC(0) ? frsqrt(a2) : TC(0);
  area = a2 * yv;
  nx = ax * yv;
  ny = ay * yv;
  nz = az * yv;
}

template <class TC>
struct Prim {
  TC rho, rux, ruy, ruz, rE, p, sr, ir;
};

template <class TC>
__device__ inline void make_prim(TC rho, TC rux, TC ruy, TC ruz, TC rE, TC g1, Prim<TC> &q) {
  q.rho = rho;
  q.rux = rux;
  q.ruy = ruy;
  q.ruz = ruz;
  q.rE = rE;
  const TC y = frsqrt(rho);  // sqrt(rho) = rho y, 1/rho = y^2
  q.ir = y * y;
  // syncState (foeuler.f90:201): p = (gamma-1)(rhoE - |rhou|^2 / (2 rho))
  q.p = g1 * (rE - TC(0.5) * (rux * rux + ruy * ruy + ruz * ruz) * q.ir);
  q.sr = rho * y;
}

// Ghost state of boundary pair b from the interior state q (foeuler.f90:214-329).
template <class TC, class TP>
__device__ inline void ghost_state(int type, const TP *par, TC nx, TC ny, TC nz, const Prim<TC> &q,
                                   TC gamm, TC r, TC &gr, TC &gux, TC &guy, TC &guz, TC &gE) {
  const TC g1 = gamm - TC(1);
  switch (type) {
    case 10: case 110: {  // static inflow (:229-249)
      TC pG = par[0], TG = par[1], ux = par[2], uy = par[3], uz = par[4];
      TC mach = dsqrt(ux * ux + uy * uy + uz * uz) / dsqrt(gamm * r * TG);
      TC rG = pG / r / TG;
      if (mach < TC(1)) {  // interior momentum over the ghost density (:245)
        ux = q.rux / rG;
        uy = q.ruy / rG;
        uz = q.ruz / rG;
      }
      gr = rG;
      gux = rG * ux; guy = rG * uy; guz = rG * uz;
      gE = rG * (TC(1) / g1 * r * TG + TC(0.5) * (ux * ux + uy * uy + uz * uz));
      return;
    }
    case 11: case 111: {  // total inflow (:250-276)
      TC pt = par[0], Tt = par[1], ux = par[2], uy = par[3], uz = par[4];
      TC uu = ux * ux + uy * uy + uz * uz;
      TC TG = Tt - TC(0.5) * g1 / gamm / r * uu;
      TC mach = dsqrt(uu) / dsqrt(gamm * r * TG);
      TC pG = pt * dpow(TC(1) + TC(0.5) * g1 * (mach * mach), -gamm / g1);
      TC rG = pG / r / TG;
      if (mach < TC(1)) {
        TC s = (q.rux * ux + q.ruy * uy + q.ruz * uz) * q.ir / uu;  // (u_i . u_G) / |u_G|^2
        ux *= s; uy *= s; uz *= s;
        uu = ux * ux + uy * uy + uz * uz;
        TG = Tt - TC(0.5) * g1 / gamm / r * uu;
        mach = dsqrt(uu) / dsqrt(gamm * r * TG);
        pG = pt * dpow(TC(1) + TC(0.5) * g1 * (mach * mach), -gamm / g1);
        rG = pG / r / TG;
      }
      gr = rG;
      gux = rG * ux; guy = rG * uy; guz = rG * uz;
      gE = rG * (TC(1) / g1 * r * TG + TC(0.5) * uu);
      return;
    }
    case 20: case 120: {  // outflow (:277-287)
      gr = q.rho;
      gux = q.rux; guy = q.ruy; guz = q.ruz;
      gE = TC(1) / g1 * par[0] + TC(0.5) * (q.rux * q.rux + q.ruy * q.ruy + q.ruz * q.ruz) * q.ir;
      return;
    }
    case 30: case 130: {  // far field by Riemann invariants (:288-325)
      TC pG = par[0], TG = par[1], ux = par[2], uy = par[3], uz = par[4];
      TC charO = ux * nx + uy * ny + uz * nz - TC(2) * dsqrt(gamm * r * TG) / g1;
      TC uix = q.rux * q.ir, uiy = q.ruy * q.ir, uiz = q.ruz * q.ir;
      TC eint = q.rE - TC(0.5) * (uix * uix + uiy * uiy + uiz * uiz) * q.rho;
      TC charI = uix * nx + uiy * ny + uiz * nz + TC(2) * dsqrt(gamm * q.ir / g1 * eint);
      TC h = TC(0.5) * (charI + charO);
      TC unx = nx * h, uny = ny * h, unz = nz * h;
      TC a = TC(0.25) * g1 * (charI - charO);
      TC utx, uty, utz, pt;
      if (unx * nx + uny * ny + unz * nz > TC(0)) {
        TC nu = nx * uix + ny * uiy + nz * uiz;
        utx = uix - nx * nu; uty = uiy - ny * nu; utz = uiz - nz * nu;
        TC mach = dsqrt(uix * uix + uiy * uiy + uiz * uiz) / dsqrt(gamm * q.ir * g1 * eint);
        pt = g1 * eint * dpow(TC(1) + TC(0.5) * g1 * (mach * mach), gamm / g1);
      } else {
        utx = ux - unx; uty = uy - uny; utz = uz - unz;
        TC mach = dsqrt(ux * ux + uy * uy + uz * uz) / dsqrt(gamm * r * TG);
        pt = pG * dpow(TC(1) + TC(0.5) * g1 * (mach * mach), gamm / g1);
      }
      ux = unx + utx; uy = uny + uty; uz = unz + utz;
      TC uu = ux * ux + uy * uy + uz * uz;
      TG = a * a / gamm / r;
      TC mach = dsqrt(uu) / a;
      pG = pt * dpow(TC(1) + TC(0.5) * g1 * (mach * mach), -gamm / g1);
      TC rG = pG / r / TG;
      gr = rG;
      gux = rG * ux; guy = rG * uy; guz = rG * uz;
      gE = rG * (TC(1) / g1 * r * TG + TC(0.5) * uu);
      return;
    }
    default: {  // wall, also for unassigned groups (:217-228)
      TC dn = TC(2) * (q.rux * nx + q.ruy * ny + q.ruz * nz);
      gr = q.rho;
      gux = q.rux - dn * nx; guy = q.ruy - dn * ny; guz = q.ruz - dn * nz;
      gE = q.rE;
      return;
    }
  }
}

// Roe flux-difference splitting across a face with unit normal n (from A to B) and area a
// (Euler.f90:90-126): flow = a * (avg(F)n - (|l_n| w_entropy + |l_-| w_- + |l_+| w_+)/2)
template <class TC>
__device__ inline void roe_flux(const Prim<TC> &A, const Prim<TC> &B, TC nx, TC ny, TC nz, TC area,
                                TC g1, TC F[5]) {
  const TC uax = A.rux * A.ir, uay = A.ruy * A.ir, uaz = A.ruz * A.ir;
  const TC ubx = B.rux * B.ir, uby = B.ruy * B.ir, ubz = B.ruz * B.ir;
  const TC Ha = (A.rE + A.p) * A.ir, Hb = (B.rE + B.p) * B.ir;
  const TC inv = frcp(A.sr + B.sr);
  const TC ux = (A.sr * uax + B.sr * ubx) * inv;
  const TC uy = (A.sr * uay + B.sr * uby) * inv;
  const TC uz = (A.sr * uaz + B.sr * ubz) * inv;
  const TC H = (A.sr * Ha + B.sr * Hb) * inv;
  const TC rhoAvg = A.sr * B.sr;
  const TC q2 = ux * ux + uy * uy + uz * uz;
  const TC c2 = g1 * (H - TC(0.5) * q2);
  const TC yc = frsqrt(c2);  // c = c2 yc, 1/c2 = yc^2
  const TC c = c2 * yc;
  const TC ic2 = yc * yc;
  const TC un = ux * nx + uy * ny + uz * nz;
  const TC dr = B.rho - A.rho, dp = B.p - A.p;
  const TC dux = ubx - uax, duy = uby - uay, duz = ubz - uaz;
  const TC dun = dux * nx + duy * ny + duz * nz;
  // entropy + shear waves (:99-107)
  const TC aU = dabs(un);
  const TC s1 = dr - dp * ic2;
  const TC ent0 = aU * s1;
  const TC entx = aU * (s1 * ux + rhoAvg * (dux - nx * dun));
  const TC enty = aU * (s1 * uy + rhoAvg * (duy - ny * dun));
  const TC entz = aU * (s1 * uz + rhoAvg * (duz - nz * dun));
  const TC ent4 = aU * (s1 * (TC(0.5) * q2) + rhoAvg * (ux * dux + uy * duy + uz * duz - un * dun));
  // acoustic waves (:108-119)
  const TC rcd = rhoAvg * c * dun;
  const TC w1 = dabs(un - c) * (TC(0.5) * (dp - rcd) * ic2);
  const TC w2 = dabs(un + c) * (TC(0.5) * (dp + rcd) * ic2);
  const TC ws = w1 + w2, wd = w2 - w1;
  // central flux average (:120-125): F(U).n = (ru.n, ru (u.n) + p n, (rE+p)(u.n))
  const TC vna = uax * nx + uay * ny + uaz * nz, vnb = ubx * nx + uby * ny + ubz * nz;
  const TC c0 = (A.rux * nx + A.ruy * ny + A.ruz * nz) + (B.rux * nx + B.ruy * ny + B.ruz * nz);
  const TC pa = A.p + B.p;
  const TC cx = A.rux * vna + B.rux * vnb + pa * nx;
  const TC cy = A.ruy * vna + B.ruy * vnb + pa * ny;
  const TC cz = A.ruz * vna + B.ruz * vnb + pa * nz;
  const TC c4 = (A.rE + A.p) * vna + (B.rE + B.p) * vnb;
  const TC h = TC(0.5) * area;
  F[0] = h * (c0 - (ent0 + ws));
  F[1] = h * (cx - (entx + ws * ux + wd * c * nx));
  F[2] = h * (cy - (enty + ws * uy + wd * c * ny));
  F[3] = h * (cz - (entz + ws * uz + wd * c * nz));
  F[4] = h * (c4 - (ent4 + ws * H + wd * c * un));
}

}  // namespace fos
