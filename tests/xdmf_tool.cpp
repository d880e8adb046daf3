// Test helper for the drop-in's XDMF output (persiansph_amd/cpp/SPH/Domain.h).
//
//   xdmf_tool write <key> <n>     builds n particles with set field values and calls
//                                 Domain::WriteXDMF(key) (host path, no GPU), with a UserOutput
//                                 that reports Prop1 = 2 Density, Prop2 = i, Prop3 = -Mass and
//                                 OutputName[0] = "ShearRate"
//   xdmf_tool read <file.hdf5>    prints every dataset of the file as
//                                 "<name> <int|float> <count>" and one value per line
//                                 (floats as their 32-bit patterns in hex)
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "SPH/Domain.h"

static void user_out(SPH::Particle *p, double &P1, double &P2, double &P3)
{
    P1 = 2.0*p->Density;
    P2 = (double)p->ID;
    P3 = -p->Mass;
}

// the field values of particle i (the Python side of the test recomputes them)
static double val(size_t i, int field, int c) { return 0.001*(double)i + 0.1*field + 0.0137*c + 1.0/3.0; }

static herr_t list_cb(hid_t, const char *name, const H5L_info_t *, void *op)
{
    static_cast<std::vector<std::string> *>(op)->push_back(name);
    return 0;
}

int main(int argc, char **argv) try
{
    if (argc < 3) { std::fprintf(stderr, "usage: xdmf_tool write <key> <n> | read <file>\n"); return 2; }
    std::string mode = argv[1];
    if (mode == "write") {
        size_t n = (size_t)std::atol(argv[3]);
        SPH::Domain dom;
        dom.Dimension = 3;
        for (size_t i = 0; i < n; ++i) {
            dom.AddSingleParticle((int)(i % 5), Vec3_t(val(i, 0, 0), val(i, 0, 1), val(i, 0, 2)), val(i, 1, 0),
                                  val(i, 2, 0), val(i, 3, 0), i % 3 == 0);
            SPH::Particle *p = dom.Particles[i];
            for (int c = 0; c < 3; ++c) { p->v(c) = val(i, 4, c); p->a(c) = val(i, 5, c); }
            p->Pressure = val(i, 6, 0);
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    p->Sigma(r, c) = val(i, 7, 3*r + c);
                    p->Strain(r, c) = val(i, 8, 3*r + c);
                }
        }
        dom.OutputName[0] = "ShearRate";
        dom.UserOutput = &user_out;
        dom.WriteXDMF(argv[2]);
        return 0;
    }
    if (mode == "read") {
        hid_t f = H5Fopen(argv[2], H5F_ACC_RDONLY, H5P_DEFAULT);
        if (f < 0) { std::fprintf(stderr, "cannot open %s\n", argv[2]); return 1; }
        std::vector<std::string> names;
        H5Literate(f, H5_INDEX_NAME, H5_ITER_INC, NULL, list_cb, &names);
        for (const std::string &nm : names) {
            hsize_t dims[4] = {0, 0, 0, 0};
            H5T_class_t cls;
            size_t sz;
            int rank = 0;
            H5LTget_dataset_ndims(f, nm.c_str(), &rank);
            H5LTget_dataset_info(f, nm.c_str(), dims, &cls, &sz);
            size_t cnt = 1;
            for (int r = 0; r < rank; ++r) cnt *= (size_t)dims[r];
            if (cls == H5T_INTEGER) {
                std::vector<int> v(cnt);
                H5LTread_dataset_int(f, nm.c_str(), v.data());
                std::printf("%s int %zu\n", nm.c_str(), cnt);
                for (int x : v) std::printf("%d\n", x);
            } else {
                std::vector<float> v(cnt);
                H5LTread_dataset_float(f, nm.c_str(), v.data());
                std::printf("%s float %zu\n", nm.c_str(), cnt);
                for (float x : v) { uint32_t b; std::memcpy(&b, &x, 4); std::printf("%08" PRIx32 "\n", b); }
            }
        }
        H5Fclose(f);
        return 0;
    }
    return 2;
}
MECHSYS_CATCH
