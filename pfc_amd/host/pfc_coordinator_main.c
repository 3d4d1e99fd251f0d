/*
 * pfc_coordinator_main.c -- the pfc-coordinator command: the argv contract,
 * messages and files of the reference's coordinator.jar
 * (src/Coordinator.java:358-397) over the C restatement in pfc_coord.c.
 *
 *   pfc-coordinator density.n coordinates.xyz [nonhexagons.nh] lPFC lA
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pfc_coord.h"

static const char* kUsage[] = {
    "This is a tool for converting PFC density fields into atom coordinates. This tool can be "
    "applied when honeycomb structures are modeled using inverted hexagonal density fields (their "
    "maxima form a honeycomb lattice); see Sec. 4.3.1 in my thesis (available: "
    "http://urn.fi/URN:ISBN:978-952-60-8608-8).",
    "",
    "Examples of valid syntax:",
    "",
    "java -jar coordinator.jar density.n coordinates.xyz [input length scale] [output length scale]",
    "",
    "java -jar coordinator.jar density.n coordinates.xyz nonhexagons.nh [input length scale] "
    "[output length scale]",
    "",
    "The program expects the names of the input density and the output coordinate files as its "
    "first arguments. A coordinate file lists each atom's type (the program assigns all atoms with "
    "a label \"A\"), its coordinates and indices of its neighbors. The program can also write out "
    "the nonhexagonal atom rings (using the indices of the member atoms) if a filename is specified "
    "(the second example above). The input length scale is the length scale of the PFC structures, "
    "typically 4*pi/sqrt(3) ~ 7.3. The output length scale is the lattice constant in the units "
    "desired, e.g., approx. 2.5 \xc3\x85 or 0.25 nm for graphene.",
};

static void die(const char* msg) {
    fprintf(stderr, "%s\n", msg);
    exit(1);
}

static double parse_double(const char* s) {
    char* end;
    errno = 0;
    const double v = strtod(s, &end);
    while (*end == ' ' || *end == '\t' || *end == '\r' || *end == '\n') end++;
    if (end == s || *end) die("Error: could not parse a number (java.lang.NumberFormatException).");
    return v;
}

int main(int argc, char** argv) {
    const int nargs = argc - 1;
    if (nargs == 0) {
        for (size_t i = 0; i < sizeof kUsage / sizeof kUsage[0]; i++) printf("%s\n", kUsage[i]);
        return 0;
    }
    if (nargs != 4 && nargs != 5) {
        printf("Error: Invalid syntax.\n");
        return 1;
    }
    const char* input = argv[1];
    FILE* f = fopen(input, "r");
    if (!f) {
        printf("Error: Input file not found.\n");
        return 1;
    }
    int Nx, Ny;
    double dx, dy;
    char line[512];
    if (!fgets(line, sizeof line, f) || sscanf(line, "%d %d %lf %lf", &Nx, &Ny, &dx, &dy) != 4 ||
        Nx < 1 || Ny < 1)
        die("Error: invalid header (java.lang.NumberFormatException).");
    int k = 2;
    const char* cout = argv[k++];
    const char* nhout = (nargs == 5) ? argv[k++] : NULL;
    const double lPFC = parse_double(argv[k++]);
    const double lA = parse_double(argv[k++]);

    double* psi = (double*)malloc(sizeof(double) * (size_t)Nx * Ny);
    if (!psi) die("Error: out of memory.");
    for (size_t n = 0; n < (size_t)Nx * Ny; n++)
        if (fscanf(f, "%lf", &psi[n]) != 1) die("Error: invalid data (java.lang.NumberFormatException).");
    fclose(f);

    pfc_coord* c = pfc_coord_run(psi, Nx, Ny, dx, dy, lPFC, lA);
    free(psi);
    if (!c) die("Error: the conversion failed (out of memory or a position outside the bins).");
    int bad;
    const int flags = pfc_coord_flags(c, &bad);
    if (flags & 16) printf("Warning: The conversion may fail for such small system sizes.\n");
    int order[4];
    const int nm = pfc_coord_messages(c, order);
    for (int i = 0; i < nm; i++) {
        switch (order[i]) {
            case 1: printf("Warning: An atom with fewer or more than three bonds was detected.\n"); break;
            case 2: printf("Warning: An atom shared by fewer or more than three rings was detected.\n"); break;
            case 4: printf("Warning: A ring with fewer than five or more than seven atoms was detected.\n"); break;
            case 8:
                printf("Error: Mismatching numbers of atoms and neighboring rings were detected for ring %d.\n", bad);
                pfc_coord_free(c);
                return 1;
        }
    }
    if (pfc_coord_write(c, cout, nhout)) die("Error: could not write the output files.");
    pfc_coord_free(c);
    return 0;
}
