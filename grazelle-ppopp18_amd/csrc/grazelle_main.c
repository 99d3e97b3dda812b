/* grazelle_main.c -- the `grazelle` executable: main.c's entry point. */
#include "grazelle_host.h"

int main(int argc, char** argv) { return gz_main(argc, argv); }
