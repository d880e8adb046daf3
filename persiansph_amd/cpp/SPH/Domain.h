// SPH::Domain drop-in for the MI355X path of PersianSPH.
//
// The user-facing C++ API of the reference (Source/Domain.h:44-166, Source/Particle.h:29-149,
// enums Domain.h:38-40) with Domain::Solve running on the GPU through the C ABI of
// persiansph_amd.h.  Model set-up (AddBoxLength, AddBoxNo, AddSingleParticle, DelParticles,
// per-particle field writes) is host code as in the reference; Solve uploads the particles,
// runs the time loop on the device, downloads at every output time and at the end.
//
// Vec3_t / Mat3_t are the parts of blitz TinyVector<double,3> / TinyMatrix<double,3,3> that
// user code touches (element access, comma initialisation, arithmetic).  Errors surface as
// `throw new Fatal(...)`, caught by MECHSYS_CATCH in main, as in the reference.
//
// Output: WriteXDMF writes the reference's <key>.hdf5 + <key>.xmf pair (HDF5 high-level API;
// link -lhdf5_hl -lhdf5).  Inside Solve, outputs with the default UserOutput are packed to FP32
// on the GPU (psph_output_begin/end) and written by a background thread while the next steps
// run; a user UserOutput gets the full state downloaded into Particles first, as it reads them.
//
// User hooks: GeneralBefore = SPH::BodyForceHook (a constant acceleration, Domain::BodyForce, on
// every free particle, as 1-Poiseuille.cpp's UserAcc) runs on the GPU.  Any other GeneralBefore
// or GeneralAfter runs on the host at its point of every step (Domain.cpp:1767, 1770): Solve
// downloads the state into Particles, calls the hook and takes back the accelerations and
// density rates it leaves (psph_set_hooks / psph_hook_accel) -- a slow path, a full download
// per hook call.
//
// In/out-flow boundaries (BC.InOutFlow with the default InCon/OutCon) run on the GPU; particles
// the recycling creates are appended to Particles at the next download.
//
// Not available on the GPU path (Solve throws Fatal): in/out-flow boundaries with particles other
// than fluid, with the Shepard filter or with user InCon/OutCon functions; hooks that change
// fields other than a and dDensity, or the particle set.
#ifndef PERSIANSPH_AMD_DOMAIN_H
#define PERSIANSPH_AMD_DOMAIN_H

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "hdf5.h"
#include "hdf5_hl.h"

#include "persiansph_amd.h"

// ----------------------------------------------------------------------------- Fatal
class Fatal
{
public:
    Fatal(char const *fmt, ...)
    {
        char buf[2048];
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        msg_ = buf;
    }
    char const *Msg() const { return msg_.c_str(); }
private:
    std::string msg_;
};

#define MECHSYS_CATCH                                                                    \
    catch (Fatal *e) { std::printf("Fatal: %s\n", e->Msg()); delete e; return 1; }      \
    catch (char const *m) { std::printf("Fatal: %s\n", m); return 1; }                  \
    catch (std::exception &e) { std::printf("Fatal: %s\n", e.what()); return 1; }

// ----------------------------------------------------------------------------- Vec3_t / Mat3_t
template <int N>
class TinyArray
{
public:
    // comma initialisation: v = 1.0, 2.0, 3.0;
    class Init
    {
    public:
        Init(double *d, int k) : d_(d), k_(k) {}
        Init operator,(double x) { if (k_ < N) d_[k_] = x; return Init(d_, k_ + 1); }
    private:
        double *d_;
        int k_;
    };
    double d[N];
    double *data() { return d; }
    const double *data() const { return d; }
};

class Vec3_t : public TinyArray<3>
{
public:
    Vec3_t() { d[0] = d[1] = d[2] = 0.0; }
    Vec3_t(double x, double y, double z) { d[0] = x; d[1] = y; d[2] = z; }
    Init operator=(double x) { d[0] = x; d[1] = x; d[2] = x; return Init(d, 1); }
    double &operator()(int i) { return d[i]; }
    double operator()(int i) const { return d[i]; }
    double &operator[](int i) { return d[i]; }
    double operator[](int i) const { return d[i]; }
    Vec3_t &operator+=(const Vec3_t &o) { for (int k = 0; k < 3; ++k) d[k] += o.d[k]; return *this; }
    Vec3_t &operator-=(const Vec3_t &o) { for (int k = 0; k < 3; ++k) d[k] -= o.d[k]; return *this; }
    Vec3_t &operator*=(double s) { for (int k = 0; k < 3; ++k) d[k] *= s; return *this; }
    Vec3_t &operator/=(double s) { for (int k = 0; k < 3; ++k) d[k] /= s; return *this; }
};

inline Vec3_t operator+(const Vec3_t &a, const Vec3_t &b) { return Vec3_t(a(0)+b(0), a(1)+b(1), a(2)+b(2)); }
inline Vec3_t operator-(const Vec3_t &a, const Vec3_t &b) { return Vec3_t(a(0)-b(0), a(1)-b(1), a(2)-b(2)); }
inline Vec3_t operator-(const Vec3_t &a) { return Vec3_t(-a(0), -a(1), -a(2)); }
inline Vec3_t operator*(double s, const Vec3_t &a) { return Vec3_t(s*a(0), s*a(1), s*a(2)); }
inline Vec3_t operator*(const Vec3_t &a, double s) { return Vec3_t(a(0)*s, a(1)*s, a(2)*s); }
inline Vec3_t operator/(const Vec3_t &a, double s) { return Vec3_t(a(0)/s, a(1)/s, a(2)/s); }
inline double dot(const Vec3_t &a, const Vec3_t &b) { return a(0)*b(0) + a(1)*b(1) + a(2)*b(2); }
inline double norm(const Vec3_t &a) { return std::sqrt(dot(a, a)); }
inline std::ostream &operator<<(std::ostream &os, const Vec3_t &a)
{ return os << "[ " << a(0) << " " << a(1) << " " << a(2) << " ]"; }

class Mat3_t : public TinyArray<9>
{
public:
    Mat3_t() { for (int k = 0; k < 9; ++k) d[k] = 0.0; }
    Init operator=(double x) { for (int k = 0; k < 9; ++k) d[k] = x; return Init(d, 1); }
    double &operator()(int i, int j) { return d[3*i+j]; }
    double operator()(int i, int j) const { return d[3*i+j]; }
};

inline void set_to_zero(Vec3_t &v) { v = 0.0, 0.0, 0.0; }
inline void set_to_zero(Mat3_t &m) { for (int k = 0; k < 9; ++k) m.d[k] = 0.0; }

// ----------------------------------------------------------------------------- Array
template <class T>
class Array : public std::vector<T>
{
public:
    size_t Size() const { return this->size(); }
    void Push(const T &x) { this->push_back(x); }
    void Resize(size_t n) { this->resize(n); }
    void Clear() { this->clear(); }
    bool Has(const T &x) const
    {
        for (const T &y : *this) if (y == x) return true;
        return false;
    }
    // order-preserving removal of the given indices (External/array.h:726-738)
    void DelItems(const Array<int> &idx)
    {
        std::vector<char> del(this->size(), 0);
        for (int i : idx) if (i >= 0 && (size_t)i < del.size()) del[i] = 1;
        size_t k = 0;
        for (size_t i = 0; i < this->size(); ++i) if (!del[i]) (*this)[k++] = (*this)[i];
        this->resize(k);
    }
};

enum Kernels_Type { Qubic_Spline = 0, Quintic = 1, Quintic_Spline = 2 };
enum Viscosity_Eq_Type { Morris = 0, Shao = 1, Incompressible_Full = 2, Takeda = 3 };
enum Gradient_Type { Squared_density = 0, Multiplied_density = 1 };

namespace SPH {

// ----------------------------------------------------------------------------- Boundary
class Boundary
{
public:
    double inDensity = 0.0, outDensity = 0.0, allDensity = 0.0;
    Vec3_t inv, outv, allv;
    bool Periodic[3] = {false, false, false};
    int InOutFlow = 0;
    double InFlowLoc1 = 0.0, InFlowLoc2 = 0.0, InFlowLoc3 = 0.0, OutFlowLoc = 0.0;
    double cellfac = 3.9;
    int inoutcounter = 0;
    bool MassConservation = false;
};

// ----------------------------------------------------------------------------- Particle
class Particle
{
public:
    bool Shepard = false;
    size_t ShepardCounter = 0, ShepardStep = 40;
    double ZWab = 0.0, SumDen = 0.0;
    bool IsFree = true;
    size_t InOut = 0;
    bool NoSlip = false;
    int ID = 0, Material = 0;
    Vec3_t x, vb, va, v, NSv, VXSPH, a;
    size_t PresEq = 0;
    double Cs = 0.0, P0 = 0.0, Pressure = 0.0;
    double Density = 0.0, Densitya = 0.0, Densityb = 0.0, dDensity = 0.0, RefDensity = 0.0;
    double FPMassC = 1.0, Mass = 0.0;
    Mat3_t StrainRate, RotationRate;
    double ShearRate = 0.0, SBar = 0.0;
    Mat3_t ShearStress, ShearStressa, ShearStressb;
    Mat3_t Sigma, Sigmaa, Sigmab;
    Mat3_t Strain, Straina, Strainb;
    Mat3_t TIR;
    double TI = 0.0, TIn = 4.0, TIInitDist = 0.0;
    double Alpha = 0.0, Beta = 0.0, Mu = 0.0, MuRef = 0.0, T0 = 0.0, m = 300.0;
    size_t VisM = 0;
    bool LES = false;
    double CSmag = 0.17;
    double G = 0.0, K = 0.0, Sigmay = 0.0;
    size_t Fail = 0;
    double c = 0.0, phi = 0.0, psi = 0.0;
    double h = 0.0;
    // fluid-solid and soil-water interaction (Particle.h:84-101)
    double FSIPressure = 0.0, FSISumKernel = 0.0;
    Vec3_t FSINSv;
    double n = 0.0, n0 = 0.0, k = 0.0, k2 = 0.0, d = 0.0, V = 0.0, RhoF = 0.0, S = 0.0;
    size_t SeepageType = 0;
    bool IsSat = false, SatCheck = false;
    int ct = 0;
    double SumKernel = 0.0;
    bool FirstStep = true;

    // Particle::Particle, Source/Particle.cpp:25-117
    Particle(int Tag, Vec3_t const &x0, Vec3_t const &v0, double Mass0, double Density0, double h0, bool Fixed = false)
    {
        ID = Tag; x = x0; v = v0; Mass = Mass0; Density = Density0; RefDensity = Density0;
        h = h0; IsFree = !Fixed;
        V = Mass/RefDensity;
    }
};

// ----------------------------------------------------------------------------- Domain
class Domain;
inline void General(Domain &) {}
inline void BodyForceHook(Domain &dom);

// default UserOutput (Domain.cpp:28-33)
inline void OutPut(Particle *, double &Prop1, double &Prop2, double &Prop3) { Prop1 = 0.0; Prop2 = 0.0; Prop3 = 0.0; }

// default InCon, OutCon, AllCon (Domain.cpp:35-51): the constant velocity and density of the
// boundary; the GPU imposes the InCon/OutCon defaults itself (a user InCon/OutCon is refused)
inline void InFlowCon(Vec3_t &, Vec3_t &Vel, double &Den, Boundary &bdry) { Vel = bdry.inv; Den = bdry.inDensity; }
inline void OutFlowCon(Vec3_t &, Vec3_t &Vel, double &Den, Boundary &bdry) { Vel = bdry.outv; Den = bdry.outDensity; }
inline void AllFlowCon(Vec3_t &, Vec3_t &Vel, double &Den, Boundary &bdry) { Vel = bdry.allv; Den = bdry.allDensity; }

class Domain
{
public:
    typedef void (*PtDom)(Domain &dom);
    typedef void (*PtOut)(Particle *Particles, double &Prop1, double &Prop2, double &Prop3);
    typedef void (*PtVel)(Vec3_t &position, Vec3_t &Vel, double &Den, Boundary &bdry);

    Array<Particle *> Particles;
    double R = 0.0;
    double sqrt_h_a = 0.0025;
    int Dimension = 2;
    double MuMax = 0.0, CsMax = 0.0;
    Vec3_t Gravity;
    Vec3_t TRPR, BLPF, CellSize;
    int CellNo[3] = {0, 0, 0};
    double hmax = 0.0, rhomax = 0.0;
    size_t SWIType = 0;
    bool FSI = false;
    double XSPH = 0.0;
    double InitialDist = 0.0;
    size_t Nproc = 1;
    Boundary BC;
    Vec3_t DomMax, DomMin;
    PtDom GeneralBefore = &General;
    PtDom GeneralAfter = &General;
    PtOut UserOutput = &OutPut;
    Vec3_t BodyForce;                    // acceleration added by BodyForceHook (this build's addition)
    PtVel InCon = &InFlowCon;
    PtVel OutCon = &OutFlowCon;
    PtVel AllCon = &AllFlowCon;
    std::string OutputName[3] = {"Property1", "Property2", "Property3"};
    size_t Scheme = 0;
    int Precision = PSPH_FP64;           // GPU arithmetic tier (this build's addition)

    Domain()
    {
        DomMax = -100000000000.0;
        DomMin = 100000000000.0;
    }
    ~Domain() { for (Particle *p : Particles) delete p; }

    void Kernel_Set(Kernels_Type const &KT) { KernelType = KT; Cellfac = (KT == 2) ? 3.0 : 2.0; }
    void Viscosity_Eq_Set(Viscosity_Eq_Type const &VQ) { VisEq = VQ; }
    void Gradient_Approach_Set(Gradient_Type const &GT) { GradientType = GT; }

    void AddSingleParticle(int tag, Vec3_t const &x, double Mass, double Density, double h, bool Fixed)
    {
        Particles.Push(new Particle(tag, x, Vec3_t(0, 0, 0), Mass, Density, h, Fixed));
    }

    // AddBoxLength, Source/Domain.cpp:294-495 (random jitter with srand(100)/rand() as there)
    void AddBoxLength(int tag, Vec3_t const &V, double Lx, double Ly, double Lz, double r, double Density,
                      double h, int type, int rotation, bool random, bool Fixed)
    {
        if (!(type == 0 || type == 1)) throw new Fatal("Packing Type is out of range (0 hexagonal, 1 cubic)");
        if (!(rotation == 0 || rotation == 90)) throw new Fatal("Packing Rotation Angle is out of range (0 or 90)");
        size_t PrePS = Particles.Size();
        double qin = 0.03;
        std::srand(100);
        auto jit = [&](double c) { return random ? c + qin*r*double(std::rand())/RAND_MAX : c; };
        if (Dimension == 3) {
            size_t k = 0;
            double zp = V(2);
            if (type == 0) {
                while (zp <= (V(2)+Lz-r)) {
                    size_t j = 0; double yp = V(1);
                    while (yp <= (V(1)+Ly-r)) {
                        size_t i = 0; double xp = V(0);
                        while (xp <= (V(0)+Lx-r)) {
                            double x, y, z;
                            if ((k%2 != 0) && (j%2 != 0)) x = V(0) + (2*i+(j%2)+(k%2)-1)*r; else x = V(0) + (2*i+(j%2)+(k%2)+1)*r;
                            y = V(1) + (std::sqrt(3.0)*(j+(1.0/3.0)*(k%2))+1)*r;
                            z = V(2) + ((2*std::sqrt(6.0)/3)*k+1)*r;
                            double xx = jit(x), yy = jit(y), zz = jit(z);
                            Particles.Push(new Particle(tag, Vec3_t(xx, yy, zz), Vec3_t(0, 0, 0), 0.0, Density, h, Fixed));
                            i++;
                            if ((k%2 != 0) && (j%2 != 0)) xp = V(0) + (2*i+(j%2)+(k%2)-1)*r; else xp = V(0) + (2*i+(j%2)+(k%2)+1)*r;
                        }
                        j++;
                        yp = V(1) + (std::sqrt(3.0)*(j+(1.0/3.0)*(k%2))+1)*r;
                    }
                    k++;
                    zp = V(2) + ((2*std::sqrt(6.0)/3)*k+1)*r;
                }
            } else {
                while (zp <= (V(2)+Lz-r)) {
                    size_t j = 0; double yp = V(1);
                    while (yp <= (V(1)+Ly-r)) {
                        size_t i = 0; double xp = V(0);
                        while (xp <= (V(0)+Lx-r)) {
                            double x = V(0) + (2.0*i+1)*r, y = V(1) + (2.0*j+1)*r, z = V(2) + (2.0*k+1)*r;
                            double xx = jit(x), yy = jit(y), zz = jit(z);
                            Particles.Push(new Particle(tag, Vec3_t(xx, yy, zz), Vec3_t(0, 0, 0), 0.0, Density, h, Fixed));
                            i++;
                            xp = V(0) + (2*i+1)*r;
                        }
                        j++;
                        yp = V(1) + (2.0*j+1)*r;
                    }
                    k++;
                    zp = V(2) + (2.0*k+1)*r;
                }
            }
            // particle mass in 3-D (Domain.cpp:401-418)
            Vec3_t Max = V;
            for (size_t i = PrePS; i < Particles.Size(); i++)
                for (int c = 0; c < 3; ++c) if (Particles[i]->x(c) > Max(c)) Max(c) = Particles[i]->x(c);
            for (int c = 0; c < 3; ++c) Max(c) += r;
            Vec3_t temp = Max - V;
            double Mass = temp(0)*temp(1)*temp(2)*Density/(Particles.Size()-PrePS);
            for (size_t i = PrePS; i < Particles.Size(); i++) Particles[i]->Mass = Mass;
        } else if (Dimension == 2) {
            if (type == 0 && rotation == 0) {
                size_t j = 0; double yp = V(1);
                while (yp <= (V(1)+Ly-r)) {
                    size_t i = 0; double xp = V(0);
                    while (xp <= (V(0)+Lx-r)) {
                        double x = V(0) + (2*i+(j%2)+1)*r, y = V(1) + (std::sqrt(3.0)*j+1)*r;
                        double xx = jit(x), yy = jit(y);
                        Particles.Push(new Particle(tag, Vec3_t(xx, yy, 0.0), Vec3_t(0, 0, 0), (std::sqrt(3.0)*r*r)*Density, Density, h, Fixed));
                        i++;
                        xp = V(0) + (2*i+(j%2)+1)*r;
                    }
                    j++;
                    yp = V(1) + (std::sqrt(3.0)*j+1)*r;
                }
            } else if (type == 0) {
                size_t i = 0; double xp = V(0);
                while (xp <= (V(0)+Lx-r)) {
                    size_t j = 0; double yp = V(1);
                    while (yp <= (V(1)+Ly-r)) {
                        double x = V(0) + (std::sqrt(3.0)*i+1)*r, y = V(1) + (2*j+(i%2)+1)*r;
                        double xx = jit(x), yy = jit(y);
                        Particles.Push(new Particle(tag, Vec3_t(xx, yy, 0.0), Vec3_t(0, 0, 0), (std::sqrt(3.0)*r*r)*Density, Density, h, Fixed));
                        j++;
                        yp = V(1) + (2*j+(i%2)+1)*r;
                    }
                    i++;
                    xp = V(0) + (std::sqrt(3.0)*i+1)*r;
                }
            } else {
                size_t j = 0; double yp = V(1);
                while (yp <= (V(1)+Ly-r)) {
                    size_t i = 0; double xp = V(0);
                    while (xp <= (V(0)+Lx-r)) {
                        double x = V(0) + (2*i+1)*r, y = V(1) + (2*j+1)*r;
                        // random cubic boxes get the hexagonal mass, as the reference does (Domain.cpp:482)
                        if (random)
                            Particles.Push(new Particle(tag, Vec3_t(jit(x), jit(y), 0.0), Vec3_t(0, 0, 0), (std::sqrt(3.0)*r*r)*Density, Density, h, Fixed));
                        else
                            Particles.Push(new Particle(tag, Vec3_t(x, y, 0.0), Vec3_t(0, 0, 0), 2.0*r*2.0*r*Density, Density, h, Fixed));
                        i++;
                        xp = V(0) + (2*i+1)*r;
                    }
                    j++;
                    yp = V(1) + (2*j+1)*r;
                }
            }
        }
        R = r;
    }

    // DelParticles, Domain.cpp:497-515
    void DelParticles(int const &Tags)
    {
        Array<int> idxs;
        for (size_t i = 0; i < Particles.Size(); ++i) if (Particles[i]->ID == Tags) idxs.Push((int)i);
        if (idxs.Size() < 1) throw new Fatal("Domain::DelParticles: Could not find any particles to delete");
        for (int i : idxs) delete Particles[i];
        Particles.DelItems(idxs);
        std::cout << "\nParticle(s) with Tag No. " << Tags << " has been deleted" << std::endl;
    }

    // Solve, Domain.cpp:1735-1797, with the time loop on the GPU
    void Solve(double tf, double dt, double dtOut, char const *TheFileKey, size_t maxidx)
    {
        std::cout << "\n--------------Solving---------------------------------------------------------------" << std::endl;
        // hooks: SPH::BodyForceHook runs on the GPU; any other GeneralBefore/GeneralAfter runs on
        // the host at its point of every step (Domain.cpp:1767, 1770), with the state downloaded
        // into Particles and the accelerations and density rates it leaves taken back
        const bool body_hook = GeneralBefore == &BodyForceHook;
        const bool hook_before = GeneralBefore != &General && !body_hook;
        const bool hook_after = GeneralAfter != &General;
        if (BC.InOutFlow > 0 && (InCon != &InFlowCon || OutCon != &OutFlowCon))
            throw new Fatal("Domain::Solve: user InCon/OutCon functions are not available on the GPU path "
                            "(it imposes the constant BC.inv/inDensity and BC.outv/outDensity of the defaults)");
        if (Particles.Size() == 0) throw new Fatal("Domain::Solve: no particles");
        psph_config cfg;
        std::memset(&cfg, 0, sizeof cfg);
        cfg.dimension = Dimension; cfg.kernel_type = (int)KernelType; cfg.visc_eq = (int)VisEq;
        cfg.gradient_type = (int)GradientType; cfg.scheme = (int)Scheme; cfg.precision = Precision;
        for (int c = 0; c < 3; ++c) {
            cfg.periodic[c] = BC.Periodic[c] ? 1 : 0;
            cfg.gravity[c] = Gravity(c); cfg.dom_max[c] = DomMax(c); cfg.dom_min[c] = DomMin(c);
            cfg.trpr[c] = TRPR(c); cfg.blpf[c] = BLPF(c);
        }
        cfg.in_out_flow = BC.InOutFlow; cfg.swi_type = (int)SWIType; cfg.fsi = FSI ? 1 : 0;
        for (int c = 0; c < 3; ++c) { cfg.in_v[c] = BC.inv(c); cfg.out_v[c] = BC.outv(c); }
        cfg.in_density = BC.inDensity; cfg.out_density = BC.outDensity; cfg.bc_cellfac = BC.cellfac;
        cfg.initial_dist = InitialDist; cfg.mass_conservation = BC.MassConservation ? 1 : 0;
        cfg.cellfac = Cellfac; cfg.xsph = XSPH; cfg.sqrt_h_a = sqrt_h_a; cfg.R = R;
        cfg.mu_max = MuMax; cfg.cs_max = CsMax;
        WholeVelocity();           // before the upload: its EOS pressure is the upload's (Domain.cpp:1656)
        std::vector<psph_particle> buf(Particles.Size());
        for (size_t i = 0; i < Particles.Size(); ++i) to_abi(*Particles[i], buf[i]);
        up_.assign(Particles.begin(), Particles.end());
        alive_.assign(up_.size(), 1);
        psph_handle *h = nullptr;
        check(psph_create(&cfg, &h));
        struct Guard {
            psph_handle *h; Domain *d;
            ~Guard() { d->join_writer(false); psph_destroy(h); }
        } guard{h, this};
        check(psph_upload(h, buf.data(), buf.size(), Time, dt));
        if (body_hook) {
            double b[3] = {BodyForce(0), BodyForce(1), BodyForce(2)};
            check(psph_set_body_force(h, b));
        }
        if (hook_before || hook_after) check(psph_set_hooks(h, hook_before ? 1 : 0, hook_after ? 1 : 0));
        psph_grid g;
        check(psph_get_grid(h, &g));
        for (int c = 0; c < 3; ++c) { TRPR(c) = g.trpr[c]; BLPF(c) = g.blpf[c]; CellSize(c) = g.cell_size[c]; CellNo[c] = g.cell_no[c]; }
        hmax = g.hmax; rhomax = g.rhomax; MuMax = g.mu_max; CsMax = g.cs_max;
        deltat = deltatint = deltatmin = dt;
        size_t idx_out = 1;
        double tout = Time;
        if (TheFileKey != NULL) {
            PrintInput(TheFileKey);
            char fn[4096];
            std::snprintf(fn, sizeof fn, "%s_Initial", TheFileKey);
            output(h, fn);
            std::cout << "\nInitial Condition has been generated\n" << std::endl;
        }
        psph_status st;
        for (;;) {
            if (!(Time < tf && idx_out <= maxidx)) break;
            check(psph_run(h, INT64_MAX, tf, tout, &st));
            Time = st.time; deltat = st.deltat; deltatmin = st.deltatmin;
            if (st.stop_reason == PSPH_STOP_HOOK_BEFORE) { run_hook(h, GeneralBefore, "GeneralBefore"); continue; }
            if (st.stop_reason == PSPH_STOP_HOOK_AFTER) { run_hook(h, GeneralAfter, "GeneralAfter"); continue; }
            if (st.stop_reason != PSPH_STOP_OUTPUT) break;
            if (hook_after) run_hook(h, GeneralAfter, "GeneralAfter");     // before the output (:1770-1773)
            if (TheFileKey != NULL) {
                char fn[4096];
                std::snprintf(fn, sizeof fn, "%s_%04zu", TheFileKey, idx_out);
                output(h, fn);
                std::cout << "\nOutput No. " << idx_out << " at " << Time << " has been generated" << std::endl;
                std::cout << "Current Time Step = " << deltat << std::endl;
            }
            idx_out++;
            tout += dtOut;
            if (idx_out > maxidx) {                      // the current step still completes
                check(psph_run(h, 1, tf, tout, &st));
                Time = st.time; deltat = st.deltat; deltatmin = st.deltatmin;
                break;
            }
        }
        join_writer(true);
        download(h);
        std::cout << "\n--------------Solving is finished---------------------------------------------------" << std::endl;
    }

    // WriteXDMF, Domain.cpp:1869-2064: <FileKey>.hdf5 with the particle fields as float and
    // <FileKey>.xmf describing them, from the host Particles
    void WriteXDMF(char const *FileKey)
    {
        const size_t n = Particles.Size();
        std::vector<float> pos(3*n), vel(3*n), acc(3*n), pr(n), den(n), mass(n), sh(n), p1(n), p2(n), p3(n);
        std::vector<float> sig(6*n), eps(6*n);
        std::vector<int> tag(n);
        static const int sym[6] = {0, 1, 2, 4, 5, 8};
        for (size_t i = 0; i < n; ++i) {
            Particle *p = Particles[i];
            for (int c = 0; c < 3; ++c) {
                pos[3*i+c] = float(p->x(c)); vel[3*i+c] = float(p->v(c)); acc[3*i+c] = float(p->a(c));
            }
            pr[i] = float(p->Pressure); den[i] = float(p->Density); mass[i] = float(p->Mass);
            sh[i] = float(p->h); tag[i] = int(p->ID);
            for (int c = 0; c < 6; ++c) { sig[6*i+c] = float(p->Sigma.d[sym[c]]); eps[6*i+c] = float(p->Strain.d[sym[c]]); }
            double P1, P2, P3;
            UserOutput(p, P1, P2, P3);
            p1[i] = float(P1); p2[i] = float(P2); p3[i] = float(P3);
        }
        std::string err = write_xdmf(FileKey, n, pos.data(), vel.data(), acc.data(), tag.data(), pr.data(),
                                     den.data(), p1.data(), p2.data(), p3.data(), mass.data(), sh.data(),
                                     sig.data(), eps.data(), OutputName);
        if (!err.empty()) throw new Fatal("Domain::WriteXDMF: %s", err.c_str());
    }

    // PrintInput, Domain.cpp:1799-1867: <FileKey>_log.dat with the run's settings
    void PrintInput(char const *FileKey)
    {
        static const char *kt[3] = {"Qubic Spline\n", "Quintic\n", "Quintic Spline\n"};
        static const char *ve[4] = {"0 => Morris et al 1997\n", "1 => Shao et al 2003\n",
                                    "2 => Real viscosity for incompressible fluids\n",
                                    "3 => Takeda et al 1994 (Real viscosity for compressible fluids)\n"};
        std::ostringstream oss;
        oss << "Dimension = " << Dimension << "D\n";
        oss << "\nKernel Type = " << (KernelType < 3 ? kt[KernelType] : "");
        oss << "\nViscosity Equation = " << (VisEq < 4 ? ve[VisEq] : "");
        oss << "\nComputational domain size\n";
        oss << "Bottom Left-Corner Front = " << BLPF << " m\n";
        oss << "Top Right-Corner Rear    = " << TRPR << " m\n";
        oss << "\nMax of the smoothing lengths, h = " << hmax << " m\n";
        oss << "Cell factor in Linked List (based on kernels) = " << Cellfac << "\n";
        oss << "\nCell Size in XYZ Directions = " << CellSize << " m\n";
        oss << "No of Cells in XYZ Directions = ( " << CellNo[0] << " , " << CellNo[1] << " , " << CellNo[2] << " )\n";
        oss << "\nInitial No of Particles = " << Particles.Size() << "\n";
        oss << "\nInitial Time Step = " << deltatint << " S\n";
        oss << "\nExternal Acceleration (Gravity enabled)= " << Gravity << " m/s2\n";
        oss << "\nNo of Threads = " << Nproc << "\n";
        oss << "\nPeriodic Boundary Condition X dir= " << (BC.Periodic[0] ? "True" : "False") << "\n";
        oss << "Periodic Boundary Condition Y dir= " << (BC.Periodic[1] ? "True" : "False") << "\n";
        oss << "Periodic Boundary Condition Z dir= " << (BC.Periodic[2] ? "True" : "False") << "\n";
        std::ofstream of(std::string(FileKey) + "_log.dat", std::ios::out);
        of << oss.str();
    }

    // WholeVelocity, Domain.cpp:1637-1660 (the pressure follows from the density at upload)
    void WholeVelocity()
    {
        if (norm(BC.allv) > 0.0 || BC.allDensity > 0.0) {
            for (size_t i = 0; i < Particles.Size(); i++) {
                Vec3_t vel; double den = 0.0;
                AllCon(Particles[i]->x, vel, den, BC);
                if (Particles[i]->IsFree && norm(BC.allv) > 0.0) Particles[i]->v = vel;
                if (Particles[i]->IsFree && BC.allDensity > 0.0) Particles[i]->Density = den;
            }
        }
    }

    // Writes <key>.hdf5 and <key>.xmf as WriteXDMF lays them out (datasets /NP, Position,
    // Velocity, Acceleration, Tag, Pressure, Density, the three user properties, Mass, h, Sigma,
    // Strain; Domain.cpp:1930-2063).  Returns an error message, empty on success.
    static std::string write_xdmf(const char *key, size_t n, const float *pos, const float *vel,
                                  const float *acc, const int *tag, const float *pr, const float *den,
                                  const float *p1, const float *p2, const float *p3, const float *mass,
                                  const float *sh, const float *sig, const float *eps,
                                  const std::string *names)
    {
        const std::string fn = std::string(key) + ".hdf5";
        hid_t f = H5Fcreate(fn.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        if (f < 0) return "cannot create " + fn;
        bool ok = true;
        int np[1] = {(int)n};
        hsize_t d1[1] = {1}, d3[1] = {3*n}, dn[1] = {n}, d6[1] = {6*n};
        ok = ok && H5LTmake_dataset_int(f, "/NP", 1, d1, np) >= 0;
        ok = ok && H5LTmake_dataset_float(f, "Position", 1, d3, pos) >= 0;
        ok = ok && H5LTmake_dataset_float(f, "Velocity", 1, d3, vel) >= 0;
        ok = ok && H5LTmake_dataset_float(f, "Acceleration", 1, d3, acc) >= 0;
        ok = ok && H5LTmake_dataset_int(f, "Tag", 1, dn, tag) >= 0;
        ok = ok && H5LTmake_dataset_float(f, "Pressure", 1, dn, pr) >= 0;
        ok = ok && H5LTmake_dataset_float(f, "Density", 1, dn, den) >= 0;
        ok = ok && H5LTmake_dataset_float(f, names[0].c_str(), 1, dn, p1) >= 0;
        ok = ok && H5LTmake_dataset_float(f, names[1].c_str(), 1, dn, p2) >= 0;
        ok = ok && H5LTmake_dataset_float(f, names[2].c_str(), 1, dn, p3) >= 0;
        ok = ok && H5LTmake_dataset_float(f, "Mass", 1, dn, mass) >= 0;
        ok = ok && H5LTmake_dataset_float(f, "h", 1, dn, sh) >= 0;
        ok = ok && H5LTmake_dataset_float(f, "Sigma", 1, d6, sig) >= 0;
        ok = ok && H5LTmake_dataset_float(f, "Strain", 1, d6, eps) >= 0;
        H5Fflush(f, H5F_SCOPE_GLOBAL);
        H5Fclose(f);
        if (!ok) return "cannot write the datasets of " + fn;
        std::ostringstream oss;
        auto item = [&](const char *name, const char *type, const char *dims_suffix, const char *num, const char *prec) {
            oss << "     <Attribute Name=\"" << name << "\" AttributeType=\"" << type << "\" Center=\"Node\">\n";
            oss << "       <DataItem Dimensions=\"" << n << dims_suffix << "\" NumberType=\"" << num << "\"" << prec << " Format=\"HDF\">\n";
            oss << "        " << fn << ":/" << name << " \n";
            oss << "       </DataItem>\n";
            oss << "     </Attribute>\n";
        };
        oss << "<?xml version=\"1.0\" ?>\n";
        oss << "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n";
        oss << "<Xdmf Version=\"2.0\">\n";
        oss << " <Domain>\n";
        oss << "   <Grid Name=\"SPHCenter\" GridType=\"Uniform\">\n";
        oss << "     <Topology TopologyType=\"Polyvertex\" NumberOfElements=\"" << n << "\"/>\n";
        oss << "     <Geometry GeometryType=\"XYZ\">\n";
        oss << "       <DataItem Format=\"HDF\" NumberType=\"Float\" Precision=\"10\" Dimensions=\"" << n << " 3\" >\n";
        oss << "        " << fn << ":/Position \n";
        oss << "       </DataItem>\n";
        oss << "     </Geometry>\n";
        oss << "     <Attribute Name=\"Tag\" AttributeType=\"Scalar\" Center=\"Node\">\n";
        oss << "       <DataItem Dimensions=\"" << n << "\" NumberType=\"Int\" Format=\"HDF\">\n";
        oss << "        " << fn << ":/Tag \n";
        oss << "       </DataItem>\n";
        oss << "     </Attribute>\n";
        item("Position", "Vector", " 3", "Float", " Precision=\"10\"");
        item("Velocity", "Vector", " 3", "Float", " Precision=\"10\"");
        item("Acceleration", "Vector", " 3", "Float", " Precision=\"10\"");
        item("Density", "Scalar", "", "Float", " Precision=\"10\" ");
        item("Pressure", "Scalar", "", "Float", " Precision=\"10\" ");
        item(names[0].c_str(), "Scalar", "", "Float", " Precision=\"10\" ");
        item(names[1].c_str(), "Scalar", "", "Float", " Precision=\"10\" ");
        item(names[2].c_str(), "Scalar", "", "Float", " Precision=\"10\" ");
        item("Sigma", "Tensor6", " 6", "Float", " Precision=\"10\"");
        item("Strain", "Tensor6", " 6", "Float", " Precision=\"10\"");
        oss << "   </Grid>\n";
        oss << " </Domain>\n";
        oss << "</Xdmf>\n";
        std::ofstream of(std::string(key) + ".xmf", std::ios::out);
        of << oss.str();
        if (!of) return "cannot write " + std::string(key) + ".xmf";
        return std::string();
    }

    double SimTime() const { return Time; }

private:
    size_t VisEq = 0, KernelType = 0, GradientType = 0;
    double Cellfac = 2.0;
    double Time = 0.0, deltat = 0.0, deltatmin = 0.0, deltatint = 0.0;

    static void check(int rc)
    {
        if (rc != 0) {
            char msg[1024];
            psph_last_error(msg, sizeof msg);
            throw new Fatal("%s (code %d)", msg, rc);
        }
    }

    static void cp9(const Mat3_t &m, double *o) { for (int k = 0; k < 9; ++k) o[k] = m.d[k]; }
    static void cp9(const double *o, Mat3_t &m) { for (int k = 0; k < 9; ++k) m.d[k] = o[k]; }
    static void cp3(const Vec3_t &v, double *o) { for (int k = 0; k < 3; ++k) o[k] = v(k); }
    static void cp3(const double *o, Vec3_t &v) { for (int k = 0; k < 3; ++k) v(k) = o[k]; }

#define PSPH_FIELDS_SCALAR(X) X(Density) X(Densitya) X(Densityb) X(dDensity) X(Pressure) X(Mass) \
    X(ZWab) X(SumDen) X(SumKernel) X(Mu) X(ShearRate) X(SBar) X(h) X(Cs) X(P0) X(RefDensity)           \
    X(FPMassC) X(Alpha) X(Beta) X(MuRef) X(T0) X(m) X(CSmag) X(TI) X(TIn) X(TIInitDist) X(G) X(K)     \
    X(Sigmay) X(c) X(phi) X(psi) X(FSIPressure) X(FSISumKernel) X(n) X(n0) X(k) X(k2) X(d) X(V) X(RhoF) X(S)
#define PSPH_FIELDS_VEC(X) X(x) X(v) X(va) X(vb) X(a) X(NSv) X(VXSPH) X(FSINSv)
#define PSPH_FIELDS_MAT(X) X(StrainRate) X(RotationRate) X(ShearStress) X(ShearStressa) X(ShearStressb) \
    X(Sigma) X(Sigmaa) X(Sigmab) X(Strain) X(Straina) X(Strainb) X(TIR)

    static void to_abi(const Particle &p, psph_particle &q)
    {
        std::memset(&q, 0, sizeof q);
#define S_(f) q.f = p.f;
#define V_(f) cp3(p.f, q.f);
#define M_(f) cp9(p.f, q.f);
        PSPH_FIELDS_SCALAR(S_) PSPH_FIELDS_VEC(V_) PSPH_FIELDS_MAT(M_)
#undef S_
#undef V_
#undef M_
        q.IsFree = p.IsFree; q.NoSlip = p.NoSlip; q.Shepard = p.Shepard; q.LES = p.LES;
        q.Material = p.Material; q.PresEq = (int32_t)p.PresEq; q.VisM = (int32_t)p.VisM; q.Fail = (int32_t)p.Fail;
        q.ID = p.ID; q.ct = p.ct; q.FirstStep = p.FirstStep; q.InOut = (int32_t)p.InOut;
        q.ShepardCounter = (int32_t)p.ShepardCounter; q.ShepardStep = (int32_t)p.ShepardStep;
        q.SeepageType = (int32_t)p.SeepageType; q.IsSat = p.IsSat; q.SatCheck = p.SatCheck;
    }

    static void from_abi(const psph_particle &q, Particle &p)
    {
#define S_(f) p.f = q.f;
#define V_(f) cp3(q.f, p.f);
#define M_(f) cp9(q.f, p.f);
        PSPH_FIELDS_SCALAR(S_) PSPH_FIELDS_VEC(V_) PSPH_FIELDS_MAT(M_)
#undef S_
#undef V_
#undef M_
        p.ct = q.ct; p.FirstStep = q.FirstStep != 0; p.ShepardCounter = (size_t)q.ShepardCounter;
        p.IsSat = q.IsSat != 0; p.SatCheck = q.SatCheck != 0;
    }

    // Copies the device state back; particles deleted by the leave check are removed from
    // Particles in order (Array::DelItems semantics), particles the in/out-flow recycling created
    // (Index past the uploaded ones, numbered in creation order) are appended as new objects.
    void download(psph_handle *h)
    {
        size_t n = 0;
        check(psph_download(h, nullptr, 0, &n));
        std::vector<psph_particle> buf(n);
        check(psph_download(h, buf.data(), n, &n));
        for (size_t k = 0; k < n; ++k) {
            const size_t ix = (size_t)buf[k].Index;
            while (up_.size() <= ix) {
                up_.push_back(new Particle(0, Vec3_t(0, 0, 0), Vec3_t(0, 0, 0), 0.0, 0.0, 0.0));
                alive_.push_back(1);
            }
        }
        std::vector<char> keep(up_.size(), 0);
        Array<Particle *> next;
        for (size_t k = 0; k < n; ++k) {
            Particle *p = up_[buf[k].Index];        // Index: position at upload or creation
            keep[buf[k].Index] = 1;
            from_abi(buf[k], *p);
            next.Push(p);
        }
        for (size_t i = 0; i < up_.size(); ++i)
            if (!keep[i] && alive_[i]) { delete up_[i]; alive_[i] = 0; }
        Particles.swap(next);
    }

    // A user hook at its point inside a step: the state is downloaded into Particles, the hook
    // runs on them, and the accelerations and density rates it leaves go back to the device
    // (psph_hook_accel).  A hook that changes any other field, or the particle set, is refused:
    // the GPU path does not take those back in the middle of a step.
    void run_hook(psph_handle *h, PtDom hook, const char *name)
    {
        download(h);
        const size_t n = Particles.Size();
        std::vector<psph_particle> pre(n), post(n);
        std::vector<Particle *> ptr(n);
        for (size_t k = 0; k < n; ++k) { to_abi(*Particles[k], pre[k]); ptr[k] = Particles[k]; }
        hook(*this);
        if (Particles.Size() != n)
            throw new Fatal("Domain::Solve: the %s hook added or removed particles, which the GPU path does "
                            "not take back in the middle of a step", name);
        std::vector<double> a(3*n), dd(n);
        for (size_t k = 0; k < n; ++k) {
            if (Particles[k] != ptr[k])
                throw new Fatal("Domain::Solve: the %s hook reordered Particles, which the GPU path does not "
                                "take back in the middle of a step", name);
            to_abi(*Particles[k], post[k]);
            const char *f = changed_field(pre[k], post[k]);
            if (f)
                throw new Fatal("Domain::Solve: the %s hook changed %s of particle %zu; on the GPU path a hook may "
                                "change only a and dDensity", name, f, k);
            for (int c = 0; c < 3; ++c) a[3*k + c] = post[k].a[c];
            dd[k] = post[k].dDensity;
        }
        check(psph_hook_accel(h, a.data(), dd.data(), n));
    }

    // The first field other than a and dDensity in which two particle records differ, or null.
    static const char *changed_field(const psph_particle &x, const psph_particle &y)
    {
#define S_(f) if (std::strcmp(#f, "dDensity") != 0 && std::memcmp(&x.f, &y.f, sizeof x.f) != 0) return #f;
#define V_(f) if (std::strcmp(#f, "a") != 0 && std::memcmp(x.f, y.f, sizeof x.f) != 0) return #f;
#define M_(f) if (std::memcmp(x.f, y.f, sizeof x.f) != 0) return #f;
#define I_(f) if (x.f != y.f) return #f;
        PSPH_FIELDS_SCALAR(S_) PSPH_FIELDS_VEC(V_) PSPH_FIELDS_MAT(M_)
        I_(IsFree) I_(NoSlip) I_(Shepard) I_(LES) I_(Material) I_(PresEq) I_(VisM) I_(Fail) I_(ID) I_(ct)
        I_(FirstStep) I_(InOut) I_(ShepardCounter) I_(ShepardStep) I_(SeepageType) I_(IsSat) I_(SatCheck)
#undef S_
#undef V_
#undef M_
#undef I_
        return nullptr;
    }

    // One output inside Solve.  With the default UserOutput the fields are packed on the GPU and
    // written by a background thread (the packed buffers stay valid until the next
    // psph_output_begin, which follows the join); otherwise the state is downloaded into
    // Particles and written from there, as UserOutput reads the particles.
    void output(psph_handle *h, const char *fn)
    {
        join_writer(true);
        if (UserOutput != &OutPut) {
            download(h);
            WriteXDMF(fn);
            return;
        }
        check(psph_output_begin(h));
        psph_output_f32 o;
        check(psph_output_end(h, &o));
        std::string key(fn);
        std::string names[3] = {OutputName[0], OutputName[1], OutputName[2]};
        writer_err_.clear();
        writer_ = std::thread([this, key, o, names] {
            std::vector<float> zero(o.n, 0.0f);
            writer_err_ = write_xdmf(key.c_str(), o.n, o.position, o.velocity, o.acceleration, o.tag,
                                     o.pressure, o.density, zero.data(), zero.data(), zero.data(),
                                     o.mass, o.h, o.sigma, o.strain, names);
        });
    }

    void join_writer(bool raise)
    {
        if (writer_.joinable()) writer_.join();
        if (raise && !writer_err_.empty()) {
            std::string e;
            e.swap(writer_err_);
            throw new Fatal("Domain::WriteXDMF: %s", e.c_str());
        }
    }

    std::vector<Particle *> up_;     // Particles as uploaded by the current Solve
    std::vector<char> alive_;
    std::thread writer_;             // background XDMF writer of the current Solve
    std::string writer_err_;
};

// The body-force hook (1-Poiseuille.cpp:23-29 with a configurable constant): a += BodyForce
// for every free particle.  Solve runs it on the GPU; this host form serves direct calls.
inline void BodyForceHook(Domain &dom)
{
    for (size_t i = 0; i < dom.Particles.Size(); i++)
        if (dom.Particles[i]->IsFree) dom.Particles[i]->a = dom.Particles[i]->a + dom.BodyForce;
}

}  // namespace SPH

#endif
