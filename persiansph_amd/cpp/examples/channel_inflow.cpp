// A 2-D channel with in/out-flow boundaries through the SPH::Domain drop-in (BC.InOutFlow = 3):
// fluid filling a 0.5 x 0.12 m channel between three layers of fixed no-slip walls, the inflow
// zone at the left imposing v = (1, 0, 0) and density 1000, the outflow zone at the right
// imposing (0.8, 0, 0) scaled by InPart/OutPart (mass conservation) and density 1000.5; the
// particles that leave are recycled into the inflow, and new ones are created when more
// inflow particles pass InFlowLoc1 than leave.  Cubic spline, h = 1.3 dx, linear EOS, Cs 20.
//
//   channel_inflow [tf [dtOut [out.bin]]]
// out.bin receives n followed by x, y, z, vx of every particle (FP64).
#include <cstdio>
#include <cstdlib>

#include "SPH/Domain.h"

int main(int argc, char **argv) try
{
    double tf = argc > 1 ? std::atof(argv[1]) : 0.03;
    double dtout = argc > 2 ? std::atof(argv[2]) : 0.01;
    const char *out = argc > 3 ? argv[3] : nullptr;

    const double dx = 0.01, L = 0.5, H = 0.12, Cs = 20.0, Mu = 1.0e-3, h = 1.3*dx;
    SPH::Domain dom;
    dom.Dimension = 2;
    dom.Scheme = 0;
    dom.InitialDist = dx;
    dom.BC.InOutFlow = 3;
    dom.BC.inv = 1.0, 0.0, 0.0;
    dom.BC.inDensity = 1000.0;
    dom.BC.outv = 0.8, 0.0, 0.0;
    dom.BC.outDensity = 1000.5;
    dom.BC.MassConservation = true;

    dom.AddBoxLength(1, Vec3_t(0.0, -3.0*dx, 0.0), L + dx/10.0, H + 6.0*dx + dx/10.0, 0, dx/2.0, 1000.0, h, 1, 0, false, false);
    for (size_t a = 0; a < dom.Particles.Size(); a++) {
        SPH::Particle *p = dom.Particles[a];
        p->Cs = Cs; p->PresEq = 0; p->Mu = Mu; p->MuRef = Mu; p->Material = 1;
        const double y = p->x(1);
        if (y < 0.0 || y > H) { p->ID = 2; p->IsFree = false; p->NoSlip = true; }
        else p->v = 1.0, 0.0, 0.0;
    }

    dom.Solve(/*tf*/tf, /*dt*/0.2*h/Cs, /*dtOut*/dtout, nullptr, 1000);
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
