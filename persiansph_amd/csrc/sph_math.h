// SPH primitives as device functions.  Each follows the expression order of the reference
// function cited beside it (Source/Functions.cpp), so that with -ffp-contract=off every
// intermediate rounds as the reference's does.  T is double for the FP64 path and float for
// the FP32 pair arithmetic of the mixed path.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace psph {

template <class T> struct K_;
template <> struct K_<double> { static constexpr double PI = 3.14159265358979323846; };
template <> struct K_<float> { static constexpr float PI = 3.14159265358979323846f; };

// 1/x and 1/sqrt(x) in FP64 from the hardware estimates refined by two Newton steps (each
// within ~2 ulp of the correctly rounded value; about half the instructions of an IEEE divide).
__device__ inline double rcp_nr(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

__device__ inline double rsq_nr(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double hx = 0.5*x;
    double e = fma(-hx*y, y, 0.5);
    y = fma(y, e, y);
    e = fma(-hx*y, y, 0.5);
    return fma(y, e, y);
}

// 1/x for the pair formulas: refined reciprocal in FP64, the hardware reciprocal (1 ulp) in
// FP32.  The reference divides; the difference is a few ulp per quotient, far inside the
// force-pass bars (1e-12 FP64, 1e-5 FP32-mixed).
__device__ inline double rcpx(double x) { return rcp_nr(x); }
__device__ inline float rcpx(float x) { return __builtin_amdgcn_rcpf(x); }

// Integer powers: the reference calls pow(x, n.0) with n = 2..5.  Here they are multiplication
// chains (x2 = x*x, then x2*x, x2*x2, x2*x2*x): at most 3 roundings, a relative error of a few
// 1e-16 against the correctly rounded power, far inside the FP64 parity bar, and about an order
// of magnitude fewer instructions than the device library's pow.  Other exponents call pow.
__device__ __forceinline__ double ipow(double x, double n)
{
    const double x2 = x*x;
    if (n == 2.0) return x2;
    if (n == 3.0) return x2*x;
    if (n == 4.0) return x2*x2;
    if (n == 5.0) return x2*x2*x;
    return pow(x, n);
}
__device__ __forceinline__ float ipow(float x, float n)
{
    int k = (int)n;
    float r = 1.0f;
    for (int i = 0; i < k; ++i) r *= x;
    return r;
}

// Kernel, Functions.cpp:25-63.  C is passed precomputed when h is known per class pair.
template <class T>
__device__ __forceinline__ T kernel_w(int dim, int kt, T q, T h)
{
    const T PI = K_<T>::PI;
    T C;
    switch (kt) {
    case 0:
        C = (dim == 2) ? T(10.0)*rcpx(T(7.0)*h*h*PI) : rcpx(h*h*h*PI);
        if (q < T(1.0)) return C*(T(1.0)-(T(3.0)/T(2.0))*q*q+(T(3.0)/T(4.0))*q*q*q);
        else if (q < T(2.0)) return C*((T(1.0)/T(4.0))*(T(2.0)-q)*(T(2.0)-q)*(T(2.0)-q));
        return T(0.0);
    case 1:
        C = (dim == 2) ? T(7.0)*rcpx(T(4.0)*h*h*PI) : T(7.0)*rcpx(T(8.0)*h*h*h*PI);
        if (q < T(2.0)) return C*ipow((T(1.0)-q/T(2.0)),T(4.0))*(T(2.0)*q+T(1.0));
        return T(0.0);
    default:
        C = (dim == 2) ? T(7.0)*rcpx(T(478.0)*h*h*PI) : rcpx(T(120.0)*h*h*h*PI);
        if (q < T(1.0)) return C*(ipow((T(3.0)-q),T(5.0))-T(6.0)*ipow((T(2.0)-q),T(5.0))+T(15.0)*ipow((T(1.0)-q),T(5.0)));
        else if (q < T(2.0)) return C*(ipow((T(3.0)-q),T(5.0))-T(6.0)*ipow((T(2.0)-q),T(5.0)));
        else if (q < T(3.0)) return C*(ipow((T(3.0)-q),T(5.0)));
        return T(0.0);
    }
}

// GradKernel, Functions.cpp:65-106: (1/r) dW/dr.
template <class T>
__device__ __forceinline__ T grad_kernel(int dim, int kt, T q, T h)
{
    const T PI = K_<T>::PI;
    T C;
    switch (kt) {
    case 0:
        C = (dim == 2) ? T(10.0)*rcpx(T(7.0)*h*h*h*PI) : rcpx(h*h*h*h*PI);
        if (q == T(0.0)) return C*rcpx(h)*(T(-3.0)+(T(9.0)/T(2.0))*q);
        else if (q < T(1.0)) return C*rcpx(q*h)*(T(-3.0)*q+(T(9.0)/T(4.0))*q*q);
        else if (q < T(2.0)) return C*rcpx(q*h)*((T(-3.0)/T(4.0))*(T(2.0)-q)*(T(2.0)-q));
        return T(0.0);
    case 1:
        C = (dim == 2) ? T(7.0)*rcpx(T(4.0)*h*h*h*PI) : T(7.0)*rcpx(T(8.0)*h*h*h*h*PI);
        if (q == T(0.0)) return C*T(-5.0)*rcpx(h)*(ipow((T(1.0)-q/T(2.0)),T(3.0))-T(3.0)*q/T(2.0)*ipow((T(1.0)-q/T(2.0)),T(2.0)));
        else if (q < T(2.0)) return C*rcpx(q*h)*T(-5.0)*q*ipow((T(1.0)-q/T(2.0)),T(3.0));
        return T(0.0);
    default:
        C = (dim == 2) ? T(7.0)*rcpx(T(478.0)*h*h*h*PI) : rcpx(T(120.0)*h*h*h*h*PI);
        if (q == T(0.0)) return C*rcpx(h)*(T(20.0)*ipow((T(3.0)-q),T(3.0))-T(120.0)*ipow((T(2.0)-q),T(3.0))+T(300.0)*ipow((T(1.0)-q),T(3.0)));
        else if (q < T(1.0)) return C*rcpx(q*h)*(T(-5.0)*ipow((T(3.0)-q),T(4.0))+T(30.0)*ipow((T(2.0)-q),T(4.0))-T(75.0)*ipow((T(1.0)-q),T(4.0)));
        else if (q < T(2.0)) return C*rcpx(q*h)*(T(-5.0)*ipow((T(3.0)-q),T(4.0))+T(30.0)*ipow((T(2.0)-q),T(4.0)));
        else if (q < T(3.0)) return C*rcpx(q*h)*(T(-5.0)*ipow((T(3.0)-q),T(4.0)));
        return T(0.0);
    }
}

// LaplaceKernel, Functions.cpp:108-146.
template <class T>
__device__ __forceinline__ T laplace_kernel(int dim, int kt, T q, T h)
{
    const T PI = K_<T>::PI;
    T C;
    T Dm1 = T(dim) - T(1.0);
    switch (kt) {
    case 0:
        C = (dim == 2) ? T(10.0)/(T(7.0)*h*h*h*h*PI) : T(1.0)/(h*h*h*h*h*PI);
        if (q < T(1.0)) return C*(T(-3.0)+(T(9.0)/T(2.0))*q) + C*Dm1/q * (T(-3.0)*q+(T(9.0)/T(4.0))*q*q);
        else if (q < T(2.0)) return C*((T(3.0)/T(2.0))*(T(2.0)-q)) + C*Dm1/q * ((T(-3.0)/T(4.0))*(T(2.0)-q)*(T(2.0)-q));
        return T(0.0);
    case 1:
        C = (dim == 2) ? T(7.0)/(T(4.0)*h*h*h*h*PI) : T(7.0)/(T(8.0)*h*h*h*h*h*PI);
        if (q < T(2.0)) return C*ipow((T(1.0)-q/T(2.0)),T(2.0))*(T(10.0)*q-T(5.0)) + C*Dm1/q * T(-5.0)*q*ipow((T(1.0)-q/T(2.0)),T(3.0));
        return T(0.0);
    default:
        C = (dim == 2) ? T(7.0)/(T(478.0)*h*h*h*h*PI) : T(1.0)/(T(120.0)*h*h*h*h*h*PI);
        if (q < T(1.0)) return C*(T(20.0)*ipow((T(3.0)-q),T(3.0))-T(120.0)*ipow((T(2)-q),T(3.0))+T(300.0)*ipow((T(1)-q),T(3.0))) + C*Dm1/q * (T(-5.0)*ipow((T(3.0)-q),T(4.0))+T(30.0)*ipow((T(2.0)-q),T(4.0))-T(75.0)*ipow((T(1.0)-q),T(4.0)));
        else if (q < T(2.0)) return C*(T(20.0)*ipow((T(3.0)-q),T(3.0))-T(120.0)*ipow((T(2)-q),T(3.0))) + C*Dm1/q * (T(-5.0)*ipow((T(3.0)-q),T(4.0))+T(30.0)*ipow((T(2.0)-q),T(4.0)));
        else if (q < T(3.0)) return C*(T(20.0)*ipow((T(3.0)-q),T(3.0))) + C*Dm1/q * (T(-5.0)*ipow((T(3.0)-q),T(4.0)));
        return T(0.0);
    }
}

// SecDerivativeKernel, Functions.cpp:148-186.
template <class T>
__device__ __forceinline__ T secderiv_kernel(int dim, int kt, T q, T h)
{
    const T PI = K_<T>::PI;
    T C;
    switch (kt) {
    case 0:
        C = (dim == 2) ? T(10.0)/(T(7.0)*h*h*h*h*PI) : T(1.0)/(h*h*h*h*h*PI);
        if (q < T(1.0)) return C*(T(-3.0)+(T(9.0)/T(2.0))*q);
        else if (q < T(2.0)) return C*((T(3.0)/T(2.0))*(T(2.0)-q));
        return T(0.0);
    case 1:
        C = (dim == 2) ? T(7.0)/(T(4.0)*h*h*h*h*PI) : T(7.0)/(T(8.0)*h*h*h*h*h*PI);
        if (q < T(2.0)) return C*ipow((T(1.0)-q/T(2.0)),T(2.0))*(T(10.0)*q-T(5.0));
        return T(0.0);
    default:
        C = (dim == 2) ? T(7.0)/(T(478.0)*h*h*h*h*PI) : T(1.0)/(T(120.0)*h*h*h*h*h*PI);
        if (q < T(1.0)) return C*(T(20.0)*ipow((T(3.0)-q),T(3.0))-T(120.0)*ipow((T(2.0)-q),T(3.0))+T(300.0)*ipow((T(1.0)-q),T(3.0)));
        else if (q < T(2.0)) return C*(T(20.0)*ipow((T(3.0)-q),T(3.0))-T(120.0)*ipow((T(2.0)-q),T(3.0)));
        else if (q < T(3.0)) return C*(T(20.0)*ipow((T(3.0)-q),T(3.0)));
        return T(0.0);
    }
}

// EOS, Functions.cpp:188-212.
__device__ __forceinline__ double eos(int eq, double Cs0, double P00, double Density, double Density0)
{
    switch (eq) {
    case 0: return P00+(Cs0*Cs0)*(Density-Density0);
    case 1: return P00+(Density0*Cs0*Cs0/7.0)*(pow(Density/Density0,7.0)-1);
    default: return (Cs0*Cs0)*Density;
    }
}

// SoundSpeed, Functions.cpp:214-238.
__device__ __forceinline__ double sound_speed(int eq, double Cs0, double Density, double Density0)
{
    if (eq == 1) return sqrt((Cs0*Cs0)*pow(Density/Density0,6.0));
    return Cs0;
}

// DensitySolid, Functions.cpp:240-264.
__device__ __forceinline__ double density_solid(int eq, double Cs0, double P00, double Pressure, double Density0)
{
    switch (eq) {
    case 0: return (Pressure-P00)/(Cs0*Cs0) + Density0;
    case 1: return pow(((Pressure-P00)*(7.0/(Density0*Cs0*Cs0))+1), 1.0/7.0) * Density0;
    default: return Pressure/(Cs0*Cs0);
    }
}

// Viscous_Force, Functions.cpp:300-332 (VI returned through v0..v2).
template <class T>
__device__ __forceinline__ void viscous_force(int viseq, T &v0, T &v1, T &v2, T Mu, T di, T dj, T GK,
                                     T b0, T b1, T b2, int dim, int kt, T rij, T h,
                                     T x0, T x1, T x2, T u0, T u1, T u2)
{
    switch (viseq) {
    case 0: {
        T s = T(2.0)*Mu/(di*dj)*GK;
        v0 = s*b0; v1 = s*b1; v2 = s*b2;
        break;
    }
    case 1: {
        T s = T(8.0)*Mu/((di+dj)*(di+dj))* GK;
        v0 = s*b0; v1 = s*b1; v2 = s*b2;
        break;
    }
    case 2: {
        T s = -Mu/(di*dj)*laplace_kernel<T>(dim, kt, rij/h, h);
        v0 = s*b0; v1 = s*b1; v2 = s*b2;
        break;
    }
    default: {
        T s = -Mu/(di*dj);
        T L = laplace_kernel<T>(dim, kt, rij/h, h);
        T S2 = secderiv_kernel<T>(dim, kt, rij/h, h);
        T dvx = u0*x0 + u1*x1 + u2*x2;
        v0 = s*(L*b0 + T(1.0)/T(3.0)*(GK*u0 + dvx*x0/(rij*rij)*(-GK+S2)));
        v1 = s*(L*b1 + T(1.0)/T(3.0)*(GK*u1 + dvx*x1/(rij*rij)*(-GK+S2)));
        v2 = s*(L*b2 + T(1.0)/T(3.0)*(GK*u2 + dvx*x2/(rij*rij)*(-GK+S2)));
        break;
    }
    }
}

// Symmetric 3x3 tensors stored as 6 components (xx, xy, xz, yy, yz, zz).
template <class T>
__device__ __forceinline__ void sym_to_full(const T s[6], T M[9])
{
    M[0] = s[0]; M[1] = s[1]; M[2] = s[2];
    M[3] = s[1]; M[4] = s[3]; M[5] = s[4];
    M[6] = s[2]; M[7] = s[4]; M[8] = s[5];
}

__device__ __forceinline__ void full_to_sym(const double M[9], double s[6])
{
    s[0] = M[0]; s[1] = M[1]; s[2] = M[2]; s[3] = M[4]; s[4] = M[5]; s[5] = M[8];
}

// Mult(Mat3,Mat3), External/matvec.h:822-828.
__device__ __forceinline__ void mat_mult(const double *A, const double *B, double *M)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            M[3*i+j] = A[3*i+2]*B[6+j] + A[3*i+1]*B[3+j] + A[3*i+0]*B[j];
}

// Symmetric eigen decomposition by cyclic Jacobi rotations, eigenvalues ascending and
// eigenvectors as columns, for the artificial-stress rotation (Functions.cpp:334-348, whose
// GSL gsl_eigen_symmv call reads the lower triangle).
__device__ __forceinline__ void eig_sym3(const double m[9], double vals[3], double vecs[9])
{
    double a[3][3], v[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            a[i][j] = (i >= j) ? m[3*i+j] : m[3*j+i];
            v[i][j] = (i == j) ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 64; ++sweep) {
        if (a[0][1] == 0.0 && a[0][2] == 0.0 && a[1][2] == 0.0) break;
        for (int r = 0; r < 3; ++r) {
            int p = (r == 2) ? 1 : 0;
            int q = (r == 0) ? 1 : 2;
            double apq = a[p][q];
            if (apq == 0.0) continue;
            double g = 100.0*fabs(apq);
            if (sweep > 3 && fabs(a[p][p]) + g == fabs(a[p][p]) && fabs(a[q][q]) + g == fabs(a[q][q])) {
                a[p][q] = a[q][p] = 0.0;
                continue;
            }
            double theta = (a[q][q] - a[p][p]) / (2.0*apq);
            double t;
            if (fabs(theta) > 1.0e150) t = 0.5/theta;
            else t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta*theta + 1.0));
            double c = 1.0/sqrt(t*t + 1.0);
            double s = t*c;
            double tau = s/(1.0 + c);
            a[p][p] -= t*apq;
            a[q][q] += t*apq;
            a[p][q] = a[q][p] = 0.0;
            int o = 3 - p - q;
            double aop = a[o][p], aoq = a[o][q];
            a[o][p] = a[p][o] = aop - s*(aoq + tau*aop);
            a[o][q] = a[q][o] = aoq + s*(aop - tau*aoq);
            for (int k = 0; k < 3; ++k) {
                double vkp = v[k][p], vkq = v[k][q];
                v[k][p] = vkp - s*(vkq + tau*vkp);
                v[k][q] = vkq + s*(vkp - tau*vkq);
            }
        }
    }
    int i0 = 0, i1 = 1, i2 = 2;
    double d[3] = {a[0][0], a[1][1], a[2][2]};
    int idx[3] = {i0, i1, i2};
    for (int i = 0; i < 2; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (d[idx[j]] < d[idx[i]]) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
    for (int k = 0; k < 3; ++k) {
        vals[k] = d[idx[k]];
        for (int i = 0; i < 3; ++i) vecs[3*i+k] = v[i][idx[k]];
    }
}

// Artificial-stress tensor TIR (Domain.cpp:955-987, :1171-1204), symmetric output.
// For 2-D only the xx, xy, yy entries are written (the reference leaves the rest untouched).
__device__ __forceinline__ void tir_tensor(int dim, const double sg[6], double TI, double Density, double out[6])
{
    if (dim == 2) {
        double teta, Sigmaxx, Sigmayy, C, S;
        if ((sg[0]-sg[3]) != 0.0) teta = 0.5*atan(2.0*sg[1]/(sg[0]-sg[3]));
        else teta = 3.14159265358979323846/4.0;
        C = cos(teta);
        S = sin(teta);
        Sigmaxx = C*C*sg[0] + 2.0*C*S*sg[1] + S*S*sg[3];
        Sigmayy = S*S*sg[0] - 2.0*C*S*sg[1] + C*C*sg[3];
        if (Sigmaxx > 0) Sigmaxx = -TI * Sigmaxx/(Density*Density); else Sigmaxx = 0.0;
        if (Sigmayy > 0) Sigmayy = -TI * Sigmayy/(Density*Density); else Sigmayy = 0.0;
        out[0] = C*C*Sigmaxx + S*S*Sigmayy;
        out[3] = S*S*Sigmaxx + C*C*Sigmayy;
        out[1] = S*C*(Sigmaxx-Sigmayy);
    } else {
        double M[9], val[3], Vec[9], Val[9], VecT[9], tmp[9], R[9];
        sym_to_full(sg, M);
        eig_sym3(M, val, Vec);
        for (int q = 0; q < 9; ++q) Val[q] = 0.0;
        for (int k = 0; k < 3; ++k) {
            double l = val[k];
            if (l > 0) l = -TI * l/(Density*Density); else l = 0.0;
            Val[4*k] = l;
        }
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) VecT[3*j+i] = Vec[3*i+j];
        mat_mult(Vec, Val, tmp);
        mat_mult(tmp, VecT, R);
        full_to_sym(R, out);
    }
}

}  // namespace psph
