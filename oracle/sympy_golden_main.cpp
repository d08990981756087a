// Evaluates the reference's sympy-generated point-triangle function
// (tests/sympy_derivatives.cpp of SymX, compiled in place by oracle/Makefile) at
// the 13 inputs given on the command line and prints the 157 outputs.
// TEST INFRASTRUCTURE ONLY (writes golden vectors for tests/).
#include <cstdio>
#include <cstdlib>

void sympy_point_triangle_distance_derivatives(double* in, double* out);

int main(int argc, char** argv)
{
    if (argc != 14) return 2;
    double in[13], out[157];
    for (int i = 0; i < 13; i++) in[i] = std::atof(argv[1 + i]);
    sympy_point_triangle_distance_derivatives(in, out);
    for (int i = 0; i < 157; i++) std::printf("%.17g\n", out[i]);
    return 0;
}
