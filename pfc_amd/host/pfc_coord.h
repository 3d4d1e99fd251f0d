/*
 * pfc_coord.h -- density field to atom coordinates, the algorithm of the
 * reference's coordinator (src/Coordinator.java:29-333, with Atom.java and
 * V.java), restated in C so that the relaxed fields can be converted without
 * a JVM and the mixed form's atom positions compared with the FP64 form's.
 *
 * Rings are the local minima of the field (8-neighbour test in raster order,
 * sub-grid position by a parabola per axis); atoms sit at the centroid of
 * every triplet of rings that are pairwise within 2 lA and whose circumcircle
 * holds no other ring; atoms sharing two rings are bonded.
 */
#ifndef PFC_COORD_H
#define PFC_COORD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pfc_coord pfc_coord;

/* Runs the conversion of psi (Ny x Nx, row-major; modified in place as the
 * reference does) for grid spacings dx, dy, PFC length scale lPFC and output
 * lattice constant lA.  Returns NULL on allocation failure.  Warnings are
 * reported through pfc_coord_flags(). */
pfc_coord* pfc_coord_run(double* psi, int Nx, int Ny, double dx, double dy, double lPFC,
                         double lA);

/* Number of atoms and rings found. */
int pfc_coord_atoms(const pfc_coord* c);
int pfc_coord_rings(const pfc_coord* c);

/* Copies atom positions (2 per atom, units of lA) into xy. */
void pfc_coord_positions(const pfc_coord* c, double* xy);

/* System width and height in units of lA. */
void pfc_coord_box(const pfc_coord* c, double* w, double* h);

/* Bit flags of the reference's checks (src/Coordinator.java:297-318):
 * 1 an atom without exactly three bonds, 2 an atom in other than three rings,
 * 4 a ring with fewer than five or more than seven atoms, 8 a ring whose atom
 * and neighbour-ring counts differ (the reference exits with code 1), 16 the
 * small-system warning of the constructor.  *bad_ring gets the first ring
 * that set flag 8. */
int pfc_coord_flags(const pfc_coord* c, int* bad_ring);

/* The flags 1, 2, 4, 8 in the order the reference prints their messages;
 * returns how many (at most 4). */
int pfc_coord_messages(const pfc_coord* c, int* order);

/* Writes the coordinate file (src/Coordinator.java:321-336) and, if
 * nonhex != NULL, the non-hexagon file (:340-355).  Returns 0 or -1. */
int pfc_coord_write(const pfc_coord* c, const char* xyz, const char* nonhex);

void pfc_coord_free(pfc_coord* c);

/* Java's Double.toString(d) for finite and special values (shortest decimal
 * that reads back to d; plain notation for 1e-3 <= |d| < 1e7, else
 * d.dddE[-]n).  buf must hold 32 bytes. */
void pfc_java_double(double d, char* buf);

#ifdef __cplusplus
}
#endif

#endif
