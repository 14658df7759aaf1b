/* shotewem_main.c -- shotewem.so: exports `main` with the signature of shotewem.c:11 so that
 * ShotProfileEWEM.jl's ccall((:main, "shotewem"), Int32, (Int32, Ptr{Ptr{UInt8}}), ...) binds to
 * the MI355X library without any change on the Julia side. */
#include "../../include/seiswem.h"

int main(int argc, char** argv) { return swem_eshot_main(argc, argv); }
