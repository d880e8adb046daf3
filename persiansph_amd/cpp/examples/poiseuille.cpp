// Plane Poiseuille flow through the SPH::Domain drop-in, set up as the reference's
// 1-Poiseuille.cpp:36-80 (periodic in x, walls |y| >= 20 dx, quintic spline, Takeda viscosity,
// modified Verlet).  The reference drives the flow with GeneralBefore = UserAcc (a += 0.002 x
// on free particles).  With hook "body" (the default) the same hook is SPH::BodyForceHook with
// BodyForce = (0.002, 0, 0), which Solve applies on the GPU; with hook "user" it is the
// reference's UserAcc itself, which Solve runs on the host at its point of every step.
//
//   poiseuille [tf [dtOut [filekey [maxidx [out.bin [hook]]]]]]
// defaults: tf = 20 s, dtOut = 0.05 s, filekey "test06", maxidx 250 (the reference's call).
// out.bin receives n followed by x, y, z, vx of every particle (FP64).  filekey "-" writes no
// output files.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "SPH/Domain.h"

using std::cout;
using std::endl;

// the GeneralBefore hook of 1-Poiseuille.cpp:22-28
void UserAcc(SPH::Domain &domi)
{
    for (size_t i = 0; i < domi.Particles.Size(); i++)
        if (domi.Particles[i]->IsFree)
            domi.Particles[i]->a += Vec3_t(0.002, 0.0, 0.0);
}

int main(int argc, char **argv) try
{
    double tf = argc > 1 ? std::atof(argv[1]) : 20.0;
    double dtout = argc > 2 ? std::atof(argv[2]) : 0.05;
    const char *key = (argc > 3 && std::string(argv[3]) != "-") ? argv[3] : (argc > 3 ? nullptr : "test06");
    size_t maxidx = argc > 4 ? (size_t)std::atol(argv[4]) : 250;
    const char *out = (argc > 5 && std::string(argv[5]) != "-") ? argv[5] : nullptr;
    const bool user_hook = argc > 6 && std::string(argv[6]) == "user";

    SPH::Domain dom;
    dom.Dimension = 2;
    dom.BC.Periodic[0] = true;
    dom.Nproc = 4;
    dom.Scheme = 0;
    dom.Kernel_Set(Quintic_Spline);
    dom.Viscosity_Eq_Set(Takeda);
    dom.Gradient_Approach_Set(Squared_density);

    double yb, h, Rho, dx, t, Cs, Mu;
    Rho = 998.21;
    Mu = 1.002e-3;
    dx = 2.5e-5;
    h = dx*1.1;
    Cs = 0.07;
    t = (0.2*h/(Cs));

    cout << "Rho = " << Rho << endl;
    cout << "Mu  = " << Mu << endl;

    if (user_hook) {
        dom.GeneralBefore = &UserAcc;
    } else {
        dom.BodyForce = 0.002, 0.0, 0.0;
        dom.GeneralBefore = &SPH::BodyForceHook;
    }
    dom.InitialDist = dx;

    dom.AddBoxLength(1, Vec3_t(0.0, -24.0*dx, 0.0), 20.0*dx + dx/10.0, 48.0*dx + dx/10.0, 0, dx/2.0, Rho, h, 1, 0, false, false);

    for (size_t a = 0; a < dom.Particles.Size(); a++) {
        dom.Particles[a]->Cs = Cs;
        dom.Particles[a]->PresEq = 0;
        dom.Particles[a]->Mu = Mu;
        dom.Particles[a]->MuRef = Mu;
        dom.Particles[a]->Material = 1;
        yb = dom.Particles[a]->x(1);
        if (yb >= 20.0*dx || yb <= -20.0*dx) {
            dom.Particles[a]->ID = 2;
            dom.Particles[a]->IsFree = false;
            dom.Particles[a]->NoSlip = true;
        }
    }

    dom.Solve(/*tf*/tf, /*dt*/t, /*dtOut*/dtout, key, maxidx);
    std::printf("particles %zu time %.9e\n", dom.Particles.Size(), dom.SimTime());
    if (out) {
        FILE *f = std::fopen(out, "wb");
        size_t n = dom.Particles.Size();
        std::fwrite(&n, sizeof n, 1, f);
        for (size_t a = 0; a < n; ++a) {
            double v[4] = {dom.Particles[a]->x(0), dom.Particles[a]->x(1), dom.Particles[a]->x(2), dom.Particles[a]->v(0)};
            std::fwrite(v, sizeof v, 1, f);
        }
        std::fclose(f);
    }
    return 0;
}
MECHSYS_CATCH
