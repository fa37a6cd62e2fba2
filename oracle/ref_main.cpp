// ref_main.cpp — runs the reference's whole objective entry FDTD_CPU()
// (FDTD_CPU.cxx:669-892, compiled unmodified from the reference tree) once,
// in a fresh process, on the 162 values read from stdin, and prints the
// returned fitness as the last line of stdout with full precision.
//
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ by oracle/Makefile when
// the reference tree is present; never linked into the product.
#include <cstdio>
#include <vector>

double FDTD_CPU(const std::vector<double>& arguments);  // FDTD_CPU.h:8

int main() {
    std::vector<double> p;
    double v;
    while (std::scanf("%lf", &v) == 1) p.push_back(v);
    const double fit = FDTD_CPU(p);
    std::printf("FDTD_CPU returned %.17g\n", fit);
    return 0;
}
