/*
 * phmesh.h -- box-mesh fixture generator for the compressible hot path.
 *
 * This is the stand-in for the reference's preprocessing chain (geombc read
 * in readnblk.f:114-376, element blocking in genblk.f:57-160, BC codes from
 * genbc*.f, CSR graph in genadj.f:1-78 / asadj.f:1-59).  It produces the
 * arrays in exactly the layouts the reference keeps in memory, so they can be
 * handed unchanged to the phg_* entry points (phgpu.h) and to the CPU oracle:
 *
 *   x(numnp,3)      column-major, coordinates                  (double)
 *   ien(nel,4)      column-major, 1-based node numbers          (int)
 *   ienb(nelb,4)    column-major, 1-based; nodes 1..3 on the wall face,
 *                   ordered so (x2-x1) x (x3-x1) points out of the box
 *                   (e3bvar.f:150-178), node 4 the interior vertex
 *   iBC(nshg)       bit codes (bc3res.f:28-153): 56 = no-slip wall
 *                   (bits 3-5), +1024 = periodic slave (bit 10)
 *   iper(nshg)      1-based final periodic master (self for non-slaves)
 *   colm(nshg+1), rowp(nnz)   1-based CSR graph of the element adjacency
 *                   (genadj.f: colm are row pointers, rowp sorted columns)
 *   y(nshg,5)       column-major state in the global order (u,v,w,p,T)
 *
 * Geometry: unit box [0,1]^3 of N^3 hexes, each split into 6 tets sharing
 * the body diagonal v000-v111 (Kuhn split); tets ordered hex-major and
 * oriented so det[x1-x4, x2-x4, x3-x4] > 0 (uniformP.c:16-33 puts node 4
 * at the 1-r-s-t vertex).  Periodic in x and z, no-slip walls at y=0,1.
 * Interior nodes are jittered by a seeded U[-j h, +j h] per coordinate;
 * with grade > 1 the y spacing grows geometrically from both walls.
 * Node numbering: n = i + (N+1)*(j + (N+1)*k), x fastest.
 */
#ifndef PHMESH_H
#define PHMESH_H

#ifdef __cplusplus
extern "C" {
#endif

/* Sizes of the N^3 box: nodes (incl. periodic slaves), tets, wall tets. */
int phm_box_counts(int N, long long* nshg, long long* nel, long long* nelb);

/* Build geometry, connectivity and BC codes.  Returns 0 on success. */
int phm_box_mesh(int N, double jitter, double grade, unsigned long long seed,
                 double* x, int* ien, int* ienb, int* iBC, int* iper);
/* As phm_box_mesh, with a capped boundary layer when first > 0: the cell
 * at each wall has height first/N, each next cell grade times the last,
 * until the uniform core height that fills the half channel (N even). */
int phm_box_mesh_layer(int N, double jitter, double grade, double first,
                       unsigned long long seed, double* x, int* ien, int* ienb,
                       int* iBC, int* iper);

/* CSR graph of the element adjacency (genadj.f:1-78, asadj.f:1-59):
 * first call fills colm and returns nnz_tot, second fills sorted rowp. */
int phm_csr_count(long long nshg, long long nel, const int* ien, int* colm,
                  long long* nnz_tot);
int phm_csr_fill(long long nshg, long long nel, const int* ien,
                 const int* colm, int* rowp);

/* Synthetic state: p = 1e5(1+0.01 s), T = 300(1+0.01 s),
 * u = c0 (mach + 0.05 s), v = w = c0 0.05 s, c0 = sqrt(gamma Rgas 300),
 * s = sum of 8 seeded sine modes with integer wavenumbers in {0,1,2}^3.
 * Writes y(nshg,5) in the global order (u,v,w,p,T). */
int phm_box_state(long long nshg, const double* x, double mach,
                  unsigned long long field_seed, double Rgas, double gamma,
                  double* y);

#ifdef __cplusplus
}
#endif
#endif
