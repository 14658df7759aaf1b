/* shotwem_main.c -- shotwem.so: exports `main` with the signature of shotwem.c:12 so that
 * ShotProfileWEM.jl:347, ccall((:main, "shotwem"), Int32, (Int32, Ptr{Ptr{UInt8}}), ...),
 * binds to the MI355X library without any change on the Julia side. */
#include "../../include/seiswem.h"

int main(int argc, char** argv) { return swem_shot_main(argc, argv); }
