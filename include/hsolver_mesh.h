/*
 * hsolver_mesh.h — host mesh module in front of the C ABI: the reference's `.in`
 * input format and its single-partition cell / region typing.
 *
 * Replaces, for one rank:
 *   StructuredMesh::readMesh  mesh/mesh.cpp:1616-1893  (xgrid/ygrid/zgrid/mobs/walls/holes)
 *   CYCASMesh::readMesh       mesh/mesh.cpp:2384-2488  (solver-side grid read)
 *   initCellsType             mesh/mesh.cpp:84-114     (mobs boxes -> solid cells)
 *   getConnectedRegions       mesh/mesh.cpp:298-454    (6-connected regions, walls cut)
 * The decomposer's processor files are not needed with one rank: the solver takes the
 * grid and the cell types directly (hs_grid).
 */
#ifndef HSOLVER_MESH_H
#define HSOLVER_MESH_H

#include "hsolver.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hs_mesh hs_mesh;

int  hs_mesh_read(const char* path, hs_mesh** out);    /* HS_E_IO if missing/unparsable */
int  hs_mesh_parse(const char* text, hs_mesh** out);   /* same grammar from a string */
int  hs_mesh_info(const hs_mesh* m, int32_t* nx, int32_t* ny, int32_t* nz, int32_t* nmobs,
                  int32_t* nwalls, int32_t* nfluid_regions, int32_t* nsolid_regions);
/* points g at the mesh's own storage (valid until hs_mesh_free) */
int  hs_mesh_grid(const hs_mesh* m, hs_grid* g);
int  hs_mesh_box(const hs_mesh* m, int32_t which, int32_t* box6);
void hs_mesh_free(hs_mesh* m);

#ifdef __cplusplus
}
#endif
#endif
