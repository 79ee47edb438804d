/* Host-side C ABI (libdgkhost.so): mesh ingestion, configuration and mesh
 * generation around the DG hot path. These replace the Gmsh SDK and config
 * calls the reference makes before entering its solver:
 *   dgh_read             gmsh::open + getElementTypes/getElementsByType
 *                        (src/dgalerkin.cpp:26, src/Mesh.cpp:25-29)
 *   dgh_config_json      config::parseConfig (src/configParser.cpp:30-150)
 *   dgh_bc_groups,
 *   dgh_group_nodes      the physical-group resolution (src/configParser.cpp:117-132)
 *                        and getNodesForPhysicalGroup (src/Mesh.cpp:371)
 *   dgh_node_bc          per element-node code giving the face BC rule of
 *                        src/Mesh.cpp:364-384 when read at a face's first node
 *   dgh_gen_*            mesh generators (the reference meshes with Gmsh .geo scripts)
 * All functions return 0 on success; dgh_last_error() describes a failure. */
#ifndef DGK_HOST_H
#define DGK_HOST_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct dgh_model dgh_model;

const char* dgh_last_error(void);
dgh_model* dgh_read(const char* path);
/* names: "tag:Name;tag:Name" for the boundary physical groups.
 * rect side_tags: {ymin, xmax, ymax, xmin}; box side_tags: {xmin, xmax, ymin, ymax, zmin, zmax};
 * line side_tags: {xmin, xmax}. A tag of 0 leaves that side without a group. */
dgh_model* dgh_gen_rect(int nx, int ny, double Lx, double Ly, int p, const int* side_tags, const char* names);
dgh_model* dgh_gen_box(int nx, int ny, int nz, double Lx, double Ly, double Lz, int p, double jitter, uint64_t seed,
                       const int* block0, const int* block1, int block_tag, const int* side_tags, const char* names);
dgh_model* dgh_gen_line(int n, double L, int p, const int* side_tags, const char* names);
void dgh_free(dgh_model* m);
int dgh_write(const dgh_model* m, const char* path);
/* out4 = {dim, order, K, Np} of the element block the solver uses */
int dgh_top_info(const dgh_model* m, int64_t* out4);
/* node_xyz [K*Np*3], node_tags [K*Np], elem_tags [K]; any pointer may be NULL */
int dgh_top_arrays(const dgh_model* m, double* node_xyz, int64_t* node_tags, int64_t* elem_tags);
int dgh_config_json(const char* config_text, char* buf, int buflen);
int dgh_bc_groups(const dgh_model* m, const char* config_text, int max_groups, int* ngroups, int* tags,
                  int* reflecting, int64_t* sizes);
int dgh_group_nodes(const dgh_model* m, int tag, int64_t* out);
int dgh_node_bc(const dgh_model* m, const char* config_text, uint8_t* node_bc);
/* Reference element of the product operators: nodes [Np*3] (Gmsh order), M [Np*Np],
 * S [dim*Np*Np], Mf [Nfp*Nfp], W [dim*Np*Np] (= M^-1 S_r), LIFT [Np*Nf*Nfp], fmask [Nf*Nfp]. */
int dgh_ref_element(int dim, int p, double* nodes, double* M, double* S, double* Mf, double* W, double* LIFT,
                    int* fmask);

#ifdef __cplusplus
}
/* C++ callers holding a parsed dgh::Config */
void dgh_node_bc_resolve(const dgh_model* m, const void* config, uint8_t* node_bc);
#endif
#endif
