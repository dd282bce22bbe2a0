// Sanitizer run of the oracle (test infrastructure, built by `make sanitize`
// with AddressSanitizer and UndefinedBehaviorSanitizer). A small reach with
// a bed, banks (the y range), a hull block under the lid, actuator-disk
// forces, wall-model coefficients, an inlet and an outlet is stepped in the
// FP64, FP32 and mixed instantiations, with the boundary rows and the uref
// column reduction, as the tests drive the oracle. Any out-of-bounds access
// or undefined operation aborts the run; the state must also stay finite.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

extern "C" {
int orc_init_f64(double* f, const double* u0, int64_t N);
int orc_init_f32(float* f, const float* u0, int64_t N);
int orc_step_f64(const double* fs, double* fd, const uint8_t* solid, const double* Fx, const double* Fy,
                 const double* Fz, double* rho, double* ux, double* uy, double* uz, const double* cfw,
                 const double* uref, int nx, int ny, int nz, double tau0, double csmag2);
int orc_step_f32(const float* fs, float* fd, const uint8_t* solid, const float* Fx, const float* Fy,
                 const float* Fz, float* rho, float* ux, float* uy, float* uz, const float* cfw,
                 const float* uref, int nx, int ny, int nz, float tau0, float csmag2);
int orc_step_mixed(const float* fs, float* fd, const uint8_t* solid, const float* Fx, const float* Fy,
                   const float* Fz, float* rho, float* ux, float* uy, float* uz, const float* cfw,
                   const float* uref, int nx, int ny, int nz, float tau0, float csmag2);
int orc_bc_f64(double* fd, const uint8_t* solid, const double* uin, int nx, int ny, int nz, double rho_out);
int orc_bc_f32(float* fd, const uint8_t* solid, const float* uin, int nx, int ny, int nz, float rho_out);
int orc_bc_mixed(float* fd, const uint8_t* solid, const float* uin, int nx, int ny, int nz, float rho_out);
int orc_uref_f64(const double* ux, const double* uy, const uint8_t* solid, const uint8_t* bedwall, double* uref,
                 int nx, int ny, int nz);
int orc_uref_f32(const float* ux, const float* uy, const uint8_t* solid, const uint8_t* bedwall, float* uref,
                 int nx, int ny, int nz);
}

namespace {

constexpr int NX = 24, NY = 10, NZ = 8;
constexpr int64_t N = (int64_t)NX * NY * NZ;

int64_t cell(int x, int y, int z) { return x + (int64_t)NX * (y + (int64_t)NY * z); }

struct Case {
    std::vector<uint8_t> solid, bedwall;
    std::vector<double> Fx, cfw, uin;
    Case() : solid(N, 0), bedwall(N, 0), Fx(N, 0.0), cfw(N, 0.0), uin((size_t)NY * NZ, 0.0) {
        for (int y = 0; y < NY; ++y)
            for (int x = 0; x < NX; ++x) solid[cell(x, y, 0)] = 1;                         // bed
        for (int z = 5; z < NZ; ++z)
            for (int y = 3; y < 7; ++y)
                for (int x = 9; x < 13; ++x) solid[cell(x, y, z)] = 1;                     // hull
        for (int z = 2; z < 4; ++z)
            for (int y = 4; y < 6; ++y) Fx[cell(14, y, z)] = 2e-4;                         // disk
        for (int z = 1; z < NZ; ++z)
            for (int y = 0; y < NY; ++y)
                for (int x = 2; x < NX - 2; ++x) {
                    const int64_t c = cell(x, y, z);
                    if (solid[c]) continue;
                    const bool bed = z == 1 || y == 0 || y == NY - 1;
                    bedwall[c] = bed;
                    if (bed) cfw[c] = 3e-3;
                }
        for (int z = 1; z < NZ; ++z)
            for (int y = 0; y < NY; ++y) uin[y + (size_t)NY * z] = 0.03;
    }
};

template <class T, class Step, class Bc, class Uref, class Init>
bool run(const Case& k, Step step, Bc bc, Uref uref_fn, Init init, int steps) {
    std::vector<T> fs(19 * N), fd(19 * N), rho(N, T(1)), ux(N), uy(N), uz(N), uref(N), u0(N), Fx(N), Fy(N), Fz(N),
        cfw(N), uin(k.uin.size());
    for (int64_t c = 0; c < N; ++c) {
        Fx[c] = T(k.Fx[c]);
        cfw[c] = T(k.cfw[c]);
        const int z = int(c / ((int64_t)NX * NY)), y = int((c / NX) % NY);
        u0[c] = k.solid[c] ? T(0) : T(k.uin[y + (size_t)NY * z]);
    }
    for (size_t j = 0; j < uin.size(); ++j) uin[j] = T(k.uin[j]);
    init(fs.data(), u0.data(), N);
    for (int s = 0; s < steps; ++s) {
        if (s % 5 == 0) uref_fn(ux.data(), uy.data(), k.solid.data(), k.bedwall.data(), uref.data(), NX, NY, NZ);
        step(fs.data(), fd.data(), k.solid.data(), Fx.data(), Fy.data(), Fz.data(), rho.data(), ux.data(), uy.data(),
             uz.data(), cfw.data(), uref.data(), NX, NY, NZ, T(0.5006), T(0.17 * 0.17));
        bc(fd.data(), k.solid.data(), uin.data(), NX, NY, NZ, T(1));
        fs.swap(fd);
    }
    for (int64_t c = 0; c < N; ++c) {
        if (k.solid[c]) continue;
        for (int i = 0; i < 19; ++i)
            if (!std::isfinite((double)fs[(size_t)i * N + c])) return false;
    }
    return true;
}

}  // namespace

int main() {
    const Case k;
    const bool a = run<double>(k, orc_step_f64, orc_bc_f64, orc_uref_f64, orc_init_f64, 40);
    const bool b = run<float>(k, orc_step_f32, orc_bc_f32, orc_uref_f32, orc_init_f32, 40);
    const bool c = run<float>(k, orc_step_mixed, orc_bc_mixed, orc_uref_f32, orc_init_f32, 40);
    if (!(a && b && c)) {
        std::printf("non-finite state: f64 %d f32 %d mixed %d\n", a, b, c);
        return 1;
    }
    std::printf("sanitize ok\n");
    return 0;
}
