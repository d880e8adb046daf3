// Dam break through the SPH::Domain drop-in, written as the reference's 3-DamBreak.cpp /
// 4-DamBreak3D.cpp drivers are: lattice box, walls fixed and no-slip, fluid above the column
// deleted, hydrostatic density, Solve.
//
//   dam_break <dim> <dx> <steps> [out.bin [filekey [out_every]]]
// runs `steps` Leapfrog steps (tf = steps * dt0 upper bound) and, if out.bin is given (or "-"
// for none), writes n followed by x, y, z, Density of every particle (FP64) for comparison.
// With a file key, Solve writes <filekey>_Initial and then an XDMF output every out_every dt0
// of simulated time (<filekey>_0001, ...), as the reference's drivers do.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "SPH/Domain.h"

int main(int argc, char **argv) try
{
    int dim = argc > 1 ? std::atoi(argv[1]) : 2;
    double dx = argc > 2 ? std::atof(argv[2]) : 0.02;
    int steps = argc > 3 ? std::atoi(argv[3]) : 10;
    const char *out = (argc > 4 && std::string(argv[4]) != "-") ? argv[4] : nullptr;
    const char *key = argc > 5 ? argv[5] : nullptr;
    double every = argc > 6 ? std::atof(argv[6]) : 1.0e12;

    SPH::Domain dom;
    dom.Dimension = dim;
    dom.Nproc = 1;
    dom.Scheme = 1;
    dom.Viscosity_Eq_Set(Morris);
    dom.Kernel_Set(Qubic_Spline);
    dom.Gradient_Approach_Set(Squared_density);

    double g = 9.81, Rho = 1000.0, Mu = 1.0e-3;
    double H = (dim == 2) ? 2.0 : 0.6, L = (dim == 2) ? 1.0 : 1.2, W = 1.0;
    double TL = (dim == 2) ? 4.0 : 3.2, TH = (dim == 2) ? 3.0 : 1.0, TW = 1.0;
    double h = 1.3*dx;
    double Cs = 10.0*std::sqrt(2.0*g*2.0);
    double t = 0.2*h/Cs;
    dom.Gravity = 0.0, -g, 0.0;
    dom.InitialDist = dx;
    if (dim == 2)
        dom.AddBoxLength(1, Vec3_t(-3.0*dx, -3.0*dx, 0.0), 6.0*dx + TL + dx/10.0, 3.0*dx + TH + dx/10.0, 0, dx/2.0, Rho, h, 1, 0, false, false);
    else
        dom.AddBoxLength(1, Vec3_t(-3.0*dx, -3.0*dx, -3.0*dx), 6.0*dx + TL + dx/10.0, 3.0*dx + TH + dx/10.0, 6.0*dx + TW + dx/10.0, dx/2.0, Rho, h, 1, 0, false, false);
    for (size_t a = 0; a < dom.Particles.Size(); a++) {
        SPH::Particle *p = dom.Particles[a];
        double xb = p->x(0), yb = p->x(1), zb = p->x(2);
        p->Cs = Cs; p->PresEq = 1; p->Mu = Mu; p->MuRef = Mu; p->Material = 1; p->Shepard = true;
        p->Alpha = 0.05;
        bool wall = xb < 0.0 || yb < 0.0 || xb > TL || (dim == 3 && (zb < 0.0 || zb > TW));
        if (wall) { p->ID = 2; p->IsFree = false; p->NoSlip = true; }
        if (yb > H && p->ID == 1) p->ID = 11;
        if (xb > L && p->ID == 1) p->ID = 11;
        if (dim == 3 && zb > W && p->ID == 1) p->ID = 11;
        if (p->ID == 1) {
            p->Density = Rho*std::pow((1+7.0*g*(H-yb)/(Cs*Cs)), (1.0/7.0));
            p->Densityb = Rho*std::pow((1+7.0*g*(H-yb)/(Cs*Cs)), (1.0/7.0));
        }
    }
    dom.DelParticles(11);
    dom.Solve(/*tf*/steps*t*(1.0 - 1e-9), /*dt*/t, /*dtOut*/every*t, key, 1000);
    std::printf("particles %zu time %.9e\n", dom.Particles.Size(), dom.SimTime());
    if (out) {
        FILE *f = std::fopen(out, "wb");
        size_t n = dom.Particles.Size();
        std::fwrite(&n, sizeof n, 1, f);
        for (size_t a = 0; a < n; ++a) {
            double v[4] = {dom.Particles[a]->x(0), dom.Particles[a]->x(1), dom.Particles[a]->x(2), dom.Particles[a]->Density};
            std::fwrite(v, sizeof v, 1, f);
        }
        std::fclose(f);
    }
    return 0;
}
MECHSYS_CATCH
