/*
 * prog_launcher -- TEST INFRASTRUCTURE ONLY.
 *
 * Runs the main() of a reference program built as a shared object
 * (oracle/_ref/librajzel_cpu.so or librajzel_gpu.so, see oracle/Makefile):
 *
 *     prog_launcher <lib.so> <parameter file> <workflow file>
 *
 * The library is opened with lazy binding: the one symbol it leaves
 * unresolved (wavenumber(), which needs FFTW) is only reached with
 * SPATFILTER=1 (precond.c:23-25).
 */
#include <dlfcn.h>
#include <stdio.h>

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s <lib.so> [program arguments]\n", argv[0]);
        return 2;
    }
    void *h = dlopen(argv[1], RTLD_LAZY | RTLD_LOCAL);
    if (!h) {
        fprintf(stderr, "dlopen: %s\n", dlerror());
        return 2;
    }
    int (*prog)(int, char **) = (int (*)(int, char **))dlsym(h, "main");
    if (!prog) {
        fprintf(stderr, "dlsym(main): %s\n", dlerror());
        return 2;
    }
    argv[1] = argv[0];
    return prog(argc - 1, argv + 1);
}
