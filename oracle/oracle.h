/* CPU oracle for the DG time-stepping hot path of povanberg/DGFEM-Acoustic.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests/, the smoke entry
 * point and the CPU-baseline leg of bench.py load it. The product library
 * (libdgk) never links, calls or falls back to it.
 *
 * It restates the reference's algorithm from its source, independently of the
 * product code (own Gmsh-order node tables, own orthonormal-basis Lagrange
 * functions, own quadrature):
 *   - setup:      Mesh::Mesh (src/Mesh.cpp:20-435), precomputeMassMatrix /
 *                 getElMassMatrix (src/Mesh.cpp:440-466), eigen::inverse/solve
 *                 (src/utils.cpp:99-109)
 *   - operators:  updateFlux (src/Mesh.cpp:569-674), precomputeFlux (500-539),
 *                 getElFlux (548-557), getElStiffVector (476-489),
 *                 eigen::minus / linEq (src/utils.cpp:118-136)
 *   - stepping:   solver::numStep (src/solver.cpp:35-52) and the RK4 body of
 *                 solver::rungeKutta (src/solver.cpp:247-285), forwardEuler's
 *                 step (src/solver.cpp:139-153)
 * Two evaluation modes share one setup:
 *   ORC_LITERAL     the reference's loop nests, containers and OpenMP structure
 *                   (quadrature-point gradients, per-element inverse mass
 *                   matrices applied transposed, nested face region);
 *   ORC_FACTORIZED  the same discrete operator factorised for affine
 *                   elements (reference-element matrices x per-element scalars).
 * Pins (tests/test_oracle.py, tests/test_oracle_pins.py, tests/test_reference_element.py):
 * reference matrices against exact rationals from Gmsh's node orderings
 * (tests/golden/reference_matrices.json) and closed forms; literal = factorized to 1e-13; hashed =
 * quadratic face search; one RK4 step = the degree-4 Taylor polynomial of the assembled operator,
 * Euler = I + A; the whole P1 operator (interior, reflecting and absorbing faces, v0 = 0 and
 * v0 != 0) against an independent closed-form assembly whose absorbing flux is 1/2 A_n^+ from an
 * eigendecomposition; exact mean-flow solutions (1D d'Alembert, 2D/3D Gaussian blob) converging
 * at order >= p + 1/2; mass conservation, energy decay with reflecting, absorbing and mixed
 * walls, steady constants with and without v0; the first-node BC rule. Mutant builds
 * (ORC_ABS_COEF, ORC_V0_SIGN in oracle.cpp) are rejected by those pins.
 * Parity unpinned: the reference's interior-penalty sign for 3D order >= 2, which follows
 * Gmsh's (unavailable) quadrature point order -- ORC_SIGN_REFERENCE takes the first point of
 * the collapsed Gauss-Legendre rule used here; Gmsh's own face-node order for order >= 4
 * files written by Gmsh (the generators and reader here agree with each other).
 */
#ifndef DGK_ORACLE_H
#define DGK_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LITERAL = 0, ORC_FACTORIZED = 1 };
enum { ORC_SIGN_REFERENCE = 0, ORC_SIGN_STABLE = 1 };

typedef struct orc_ctx orc_ctx;

/* node_xyz [K*Np*3], node_tags [K*Np] in Gmsh element-node order.
 * Boundary groups: ngroups entries, ascending physical tag; group_reflecting[i]
 * is 1 for "Reflecting", 0 for "Absorbing"; group_nodes[i] holds group_sizes[i]
 * node tags (getNodesForPhysicalGroup). literal_setup bit 0 uses the reference's
 * quadratic face search (tiny meshes only); bit 1 skips the quadrature-point arrays and
 * per-element inverse mass matrices, leaving only ORC_FACTORIZED evaluation (large meshes).
 * Returns NULL on error (err filled). */
orc_ctx* orc_create(int dim, int order, int64_t K, const double* node_xyz, const int64_t* node_tags, int ngroups,
                    const int* group_reflecting, const int64_t* const* group_nodes, const int64_t* group_sizes,
                    double c0, double rho0, const double* v0, int sign_mode, int nthreads, int literal_setup,
                    char* err, int errlen);
void orc_destroy(orc_ctx*);

/* sizes: out[0]=Np out[1]=Nf out[2]=Nfp out[3]=F out[4]=F_boundary out[5]=G out[6]=Gf out[7]=fc */
void orc_sizes(const orc_ctx*, int64_t* out);

/* Connectivity as the reference stores it, flattened:
 * elFIds[K*Nf], elFOrientation[K*Nf] (after the boundary re-orientation),
 * fNbr[F*2] (-1 when absent), fNToElNIds[F*Nfp*2] (-1 when absent), fBC[F] (-1 interior). */
void orc_connectivity(const orc_ctx*, int32_t* elFIds, int32_t* elFOrientation, int32_t* fNbr, int32_t* fNToElNIds,
                      int32_t* fBC);

/* Face normal at the first face quadrature point, F*3 (as stored, boundary faces outward). */
void orc_face_normals(const orc_ctx*, double* out);

/* Per-element inverse mass matrices as stored (row-major), K*Np*Np (literal setup only). */
void orc_mass_inverse(const orc_ctx*, double* out);
int orc_has_literal(const orc_ctx*);

/* out = numStep(beta=0) after updateFlux: out = dt*Minv^T (S - F) for u, [4][K*Np].
 * orc_rhs and orc_step return 2 when ORC_LITERAL is asked without the literal setup. */
int orc_rhs(orc_ctx*, const double* u, double* out, double dt, int mode);

/* Advance u [4][K*Np] by nsteps classic RK4 steps (or forward Euler when euler=1).
 * Sources (nsrc x 9 doubles: pole,x,y,z,radius,amp,freq,phase,duration) are imposed at
 * the start of step i with time t_steps[i] (src/solver.cpp:248-256). */
int orc_step(orc_ctx*, double* u, int nsteps, double dt, const double* t_steps, int nsrc, const double* src,
             int mode, int euler);

/* Reference-element matrices of the oracle's own basis, for cross-checks:
 * M [Np*Np], S [dim*Np*Np] (S[r][i][j] = int d_r phi_i phi_j), Mf [Nfp*Nfp]. */
int orc_ref_matrices(int dim, int order, double* M, double* S, double* Mf);
/* Reference coordinates (Gmsh order) of the element nodes, Np*3. */
int orc_ref_nodes(int dim, int order, double* xyz);

#ifdef __cplusplus
}
#endif
#endif
