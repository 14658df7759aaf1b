/* compute_angles_main.c -- compute_angles.so: exports `main` with the signature of
 * compute_angles.c:12 so that ComputeAngles.jl's ccall((:main, "compute_angles"), Int32,
 * (Int32, Ptr{Ptr{UInt8}}), ...) binds to the MI355X library without any change on the Julia side. */
#include "../../include/seiswem.h"

int main(int argc, char** argv) { return swem_angles_main(argc, argv); }
