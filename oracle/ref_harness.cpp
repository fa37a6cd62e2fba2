// ref_harness.cpp — drives the reference's own CPU functions (compiled
// unmodified from FDTD_CPU.cxx and FDTD_common.cxx under the reference tree)
// so that the oracle can be checked against them function by function.
//
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ by oracle/Makefile when
// the reference tree is present; never linked into the product.
//
// The sequence below is the body of FDTD_CPU() (FDTD_CPU.cxx:669-892) for ONE
// excitation angle, truncated to `nsteps` steps, with the far-field loop's
// index initialised to 0 (FDTD_CPU.cxx:816 leaves it uninitialised).  Every
// numerical function called here is the reference's own.

#include <cstdio>
#include <cstring>
#include <vector>

using std::vector;

// Layout-compatible declaration of the reference's complex type
// (FDTD_CPU.cxx:31-55: two floats r, i).
struct cpu_cuComplex {
    float r;
    float i;
};

// Reference functions with external linkage (FDTD_CPU.cxx, FDTD_common.cxx).
void cpu_H_field_update(float*, float*, float*, float*, float*, float*, float*, float*, float*, float*);
void E_field_update(int*, float*, float*, float*, float*, float*, float*, float*, float*, float*, float*, float*,
                    float*, float*, float*, float*, float);
void calculate_JandM(float*, int*, float*, float*, float*, cpu_cuComplex*, cpu_cuComplex*, cpu_cuComplex*,
                     cpu_cuComplex*, cpu_cuComplex*, cpu_cuComplex*, cpu_cuComplex*, cpu_cuComplex*);
void N2FPostProcess(float*, float, cpu_cuComplex*, cpu_cuComplex*, cpu_cuComplex*, cpu_cuComplex*, cpu_cuComplex*,
                    cpu_cuComplex*, cpu_cuComplex*, cpu_cuComplex*, cpu_cuComplex*, cpu_cuComplex*);
void eps_r_z_init(float*, const vector<float>&);
void sigma_e_z_init(float*, float*, const vector<float>&);
void Ceze_init(float*, float*, float*);
void Cezhy_init(float*, float*, float*, float*);
void k_e_init(float*);
void k_m_init(float*);
void sigma_e_pml_init(float*);
void sigma_m_pml_init(float*, float*);
void alpha_e_init(float*);
void alpha_m_init(float*, float*);
void bex_init(float*, float*, float*, float*);
void aex_init(float*, float*, float*, float*, float*);
void bmx_init(float*, float*, float*, float*);
void amx_init(float*, float*, float*, float*, float*);
void cpu_scattered_parameter_init(float*, float*, float*, float*);
int CPUgetxfromthreadIdNF2FF(int);
int CPUgetyfromthreadIdNF2FF(int);
float calc_incident_power(float);
float fitness(float*, int, float*);  // FDTD_common.cxx:53-61

// Geometry of the compiled reference (FDTD_common.hxx:24-41 at dx = 1 mm).
static const int NX = 401, NY = 401, NC = 10, DNF = 100, NFREQ = 10, NOBS = 100;
static const int SIZE_X_SIDE = NY - 2 * DNF, SIZE_Y_SIDE = NX - 2 * DNF - 2;
static const double PI_ = 3.141592653589793238;

extern "C" {

int ref_geometry(int* out) {
    out[0] = NX; out[1] = NY; out[2] = NC; out[3] = DNF; out[4] = NFREQ; out[5] = NOBS;
    out[6] = 2 * NX - 8 * DNF + 2 * NY - 4;
    return 0;
}

int ref_contour(int* xs, int* ys) {
    int P = 2 * NX - 8 * DNF + 2 * NY - 4;
    for (int i = 0; i < P; ++i) { xs[i] = CPUgetxfromthreadIdNF2FF(i); ys[i] = CPUgetyfromthreadIdNF2FF(i); }
    return P;
}

float ref_incident_power(float f) { return calc_incident_power(f); }

// The reference's generic fitness (FDTD_common.cxx:53-61) on caller arrays.
float ref_fitness(float* D, int n, float* measurement) { return fitness(D, n, measurement); }

// CPML tables as FDTD_CPU() builds them (FDTD_CPU.cxx:765-776): 6 x 10 floats
// in the order kex, kmx, aex, bex, amx, bmx.
int ref_cpml_tables(float* out) {
    float sigma_e_pml[NC], sigma_m_pml[NC], alpha_e[NC], alpha_m[NC];
    float *kex = out, *kmx = out + NC, *aex = out + 2 * NC, *bex = out + 3 * NC, *amx = out + 4 * NC,
          *bmx = out + 5 * NC;
    k_e_init(kex);
    k_m_init(kmx);
    sigma_e_pml_init(sigma_e_pml);
    sigma_m_pml_init(sigma_m_pml, sigma_e_pml);
    alpha_e_init(alpha_e);
    alpha_m_init(alpha_e, alpha_m);
    bex_init(bex, sigma_e_pml, kex, alpha_e);
    aex_init(aex, sigma_e_pml, kex, alpha_e, bex);
    bmx_init(bmx, sigma_m_pml, kmx, alpha_m);
    amx_init(amx, sigma_m_pml, kmx, alpha_m, bmx);
    return 0;
}

// Material maps and coefficients for a 162-vector (9x9 patches of eps_r, sigma).
int ref_materials(const double* params, float* eps, float* sigma, float* Ceze, float* Cezhy, float* Cezeic,
                  float* Cezeip) {
    vector<float> image;
    for (int k = 0; k < 162; ++k) image.push_back((float)params[k]);
    float sigma_e_pml[NC];
    float kex[NC];
    k_e_init(kex);
    sigma_e_pml_init(sigma_e_pml);
    eps_r_z_init(eps, image);
    sigma_e_z_init(sigma, sigma_e_pml, image);
    Ceze_init(eps, sigma, Ceze);
    Cezhy_init(eps, sigma, Cezhy, kex);
    cpu_scattered_parameter_init(eps, sigma, Cezeic, Cezeip);
    return 0;
}

// One excitation of FDTD_CPU for `nsteps` steps.  Outputs (any may be NULL):
//   Ez (NX+1)(NY+1), Hx, Hy NX*NY, trace[nsteps] = Ez at the centre after each
//   step (the value FDTD_CPU.cxx:808 prints), cj: the 8 accumulator arrays
//   concatenated as cjzyn, cjzxp, cjzyp, cjzxn, cmxyn, cmyxp, cmxyp, cmyxn
//   (interleaved complex, [freq][leg point]), D[NFREQ*NOBS].
int ref_run(const double* params, int phi_index, int nsteps, float* Ez_out, float* Hx_out, float* Hy_out,
            float* trace, float* cj_out, float* D_out) {
    const int N1 = (NX + 1) * (NY + 1);
    vector<float> image;
    for (int k = 0; k < 162; ++k) image.push_back((float)params[k]);
    vector<float> Ceze(N1), Cezhy(N1), Cezeic(N1), Cezeip(N1), Ez(N1, 0.f), eps(N1), sig(N1);
    vector<float> Hy(NX * NY, 0.f), Hx(NX * NY, 0.f);
    float kex[NC], kmx[NC], aex[NC], bex[NC], amx[NC], bmx[NC], alpha_e[NC], alpha_m[NC], sigma_e_pml[NC],
        sigma_m_pml[NC];
    vector<float> Psi_ezy(NY * 20, 0.f), Psi_ezx(NX * 20, 0.f), Psi_hyx(NY * 20, 0.f), Psi_hxy(NX * 20, 0.f);
    const int szy = SIZE_Y_SIDE * NFREQ, szx = SIZE_X_SIDE * NFREQ;
    vector<cpu_cuComplex> cjzyp(szy), cjzyn(szy), cjzxp(szx), cjzxn(szx), cmxyn(szy), cmxyp(szy), cmyxp(szx),
        cmyxn(szx);
    for (auto* v : {&cjzyp, &cjzyn, &cjzxp, &cjzxn, &cmxyn, &cmxyp, &cmyxp, &cmyxn})
        for (auto& z : *v) z = cpu_cuComplex{0.f, 0.f};
    // FDTD_CPU.cxx:758-781 (sigma_e_pml is initialised after its first use
    // there; sigma_e_z_init ignores it)
    eps_r_z_init(eps.data(), image);
    sigma_e_z_init(sig.data(), sigma_e_pml, image);
    Ceze_init(eps.data(), sig.data(), Ceze.data());
    k_e_init(kex);
    k_m_init(kmx);
    Cezhy_init(eps.data(), sig.data(), Cezhy.data(), kex);
    sigma_e_pml_init(sigma_e_pml);
    sigma_m_pml_init(sigma_m_pml, sigma_e_pml);
    alpha_e_init(alpha_e);
    alpha_m_init(alpha_e, alpha_m);
    bex_init(bex, sigma_e_pml, kex, alpha_e);
    aex_init(aex, sigma_e_pml, kex, alpha_e, bex);
    bmx_init(bmx, sigma_m_pml, kmx, alpha_m);
    amx_init(amx, sigma_m_pml, kmx, alpha_m, bmx);
    cpu_scattered_parameter_init(eps.data(), sig.data(), Cezeic.data(), Cezeip.data());
    float Phi = (float)phi_index * 2 * PI_ / 4;  // FDTD_CPU.cxx:801
    for (int i = 0; i < nsteps; ++i) {
        cpu_H_field_update(Hy.data(), Hx.data(), Ez.data(), bmx, Psi_hyx.data(), amx, bmx, amx, Psi_hxy.data(),
                           kmx);
        E_field_update(&i, Ez.data(), Hy.data(), Hx.data(), Psi_ezx.data(), aex, aex, bex, bex, Psi_ezy.data(),
                       kex, Cezhy.data(), Cezhy.data(), Ceze.data(), Cezeip.data(), Cezeic.data(), Phi);
        if (trace) trace[i] = Ez[NX / 2 + (NY / 2) * (NX + 1)];
        for (int f = 0; f < NFREQ; ++f) {
            float freq = 1e9 + ((1e10 - 1e9) / (NFREQ - 1)) * f;
            calculate_JandM(&freq, &i, Ez.data(), Hy.data(), Hx.data(), cjzxp.data() + f * SIZE_X_SIDE,
                            cjzyp.data() + f * SIZE_Y_SIDE, cjzxn.data() + f * SIZE_X_SIDE,
                            cjzyn.data() + f * SIZE_Y_SIDE, cmxyp.data() + f * SIZE_Y_SIDE,
                            cmyxp.data() + f * SIZE_X_SIDE, cmxyn.data() + f * SIZE_Y_SIDE,
                            cmyxn.data() + f * SIZE_X_SIDE);
        }
    }
    if (Ez_out) std::memcpy(Ez_out, Ez.data(), sizeof(float) * N1);
    if (Hx_out) std::memcpy(Hx_out, Hx.data(), sizeof(float) * NX * NY);
    if (Hy_out) std::memcpy(Hy_out, Hy.data(), sizeof(float) * NX * NY);
    if (cj_out) {
        float* o = cj_out;
        for (auto* v : {&cjzyn, &cjzxp, &cjzyp, &cjzxn, &cmxyn, &cmyxp, &cmxyp, &cmyxn}) {
            std::memcpy(o, v->data(), sizeof(cpu_cuComplex) * v->size());
            o += 2 * v->size();
        }
    }
    if (D_out) {
        const int P = 2 * NX - 8 * DNF + 2 * NY - 4;
        vector<cpu_cuComplex> L(P), N(P);
        for (int f = 0; f < NFREQ; ++f) {
            for (auto& z : L) z = cpu_cuComplex{0.f, 0.f};
            for (auto& z : N) z = cpu_cuComplex{0.f, 0.f};
            float freq = 1e9 + ((1e10 - 1e9) / (NFREQ - 1)) * f;
            N2FPostProcess(D_out + f * NOBS, freq, N.data(), L.data(), cjzxp.data() + SIZE_X_SIDE * f,
                           cjzyp.data() + SIZE_Y_SIDE * f, cjzxn.data() + SIZE_X_SIDE * f,
                           cjzyn.data() + SIZE_Y_SIDE * f, cmxyp.data() + SIZE_Y_SIDE * f,
                           cmyxp.data() + SIZE_X_SIDE * f, cmxyn.data() + SIZE_Y_SIDE * f,
                           cmyxn.data() + SIZE_X_SIDE * f);
        }
    }
    return 0;
}

}  // extern "C"
