/*
 * hsolver_mesh.h — host mesh module in front of the C ABI: the reference's `.in`
 * input format and its single-partition cell / region typing.
 *
 * Replaces, for one rank:
 *   StructuredMesh::readMesh  mesh/mesh.cpp:1616-1893  (xgrid/ygrid/zgrid/mobs/walls/holes,
 *                                                       or the cylindrical nyr(1)= y form)
 *   CYCASMesh::readMesh       mesh/mesh.cpp:2384-2488  (solver-side grid read)
 *   initCellsType             mesh/mesh.cpp:84-114     (mobs boxes -> solid cells)
 *   getConnectedRegions       mesh/mesh.cpp:298-454    (6-connected regions, walls cut)
 * and the processor files of one rank (processor0/cells.txt, faces.txt, boundary.txt):
 *   decomposeMesh + writeProcFile mesh/mesh.cpp:735-1581, 2277-2374  (nParts = 1)
 *   CYCASMesh::readCells / readFaces / readBoundaries mesh/mesh.cpp:2507-2642
 * The solver itself takes the grid and the cell types (hs_grid); the processor files
 * carry the same typing for hosts that keep the decomposer's file boundary.
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
/* The cylindrical y form (no ygrid; yl(1)=a, nyr(1)=n, yl(2)=b; mesh/mesh.cpp:1644-1700,
 * perioFlag mesh/mesh.h:128): span = b - a (0 for the ygrid form), periodic = 1 when the
 * span is 360.  A periodic mesh has n cells in y (mesh.cpp:66-69), no y wall faces and
 * regions that wrap from row 1 to row n (mesh.cpp:145-165, 322-327), as the decomposer
 * types it; its grid view (hs_mesh_grid) has n + 1 y vertices with row n as wide as the
 * others, and the solver keeps walls at the y ends as the reference's solver does. */
int  hs_mesh_y_period(const hs_mesh* m, int32_t* span, int32_t* periodic);
/* points g at the mesh's own storage (valid until hs_mesh_free) */
int  hs_mesh_grid(const hs_mesh* m, hs_grid* g);
int  hs_mesh_box(const hs_mesh* m, int32_t which, int32_t* box6);

/* Writes <dir>/processor0/{cells,faces,boundary}.txt as the reference's decomposer does
 * with one part (decompose.cpp:5-31 with nParts = 1; dir NULL = the working directory,
 * which must exist).  Cells are 1-based global indices, x fastest; faces are global face
 * ids (x faces, then y, then z) with owner / neighbour positions in the cell list,
 * internal faces first, then the "ib_0_0" (fluid-solid) and "wall" groups.  After
 * hs_mesh_read_proc it writes the files that were read.  HS_E_ARG when the typing has
 * more than one fluid or more than one solid region (one part holds one of each: the
 * reference's decomposer stops, mesh.cpp:762-765, or drops solid regions), HS_E_IO when
 * a file cannot be written. */
int  hs_mesh_write_proc(const hs_mesh* m, const char* dir);
/* Reads <dir>/processor0/ as CYCASMesh::readCells/readFaces/readBoundaries do
 * (main.cpp:18-21) and takes the cell types from its fluid and mobs cell lists.
 * HS_E_IO when a file is missing or malformed, an index is out of range, a face is not
 * a face of the cells it names, or the lists do not cover every cell exactly once (the
 * files of one rank of a multi-part decomposition do not); m is unchanged then.  The
 * types are written in place, so a grid view from hs_mesh_grid sees them. */
int  hs_mesh_read_proc(hs_mesh* m, const char* dir);
/* The processor view of class which (0 = fluid, 1 = mobs): as read, or as
 * hs_mesh_write_proc would write it.  A class with no cells reports ncells = 0. */
int  hs_mesh_proc_part(const hs_mesh* m, int32_t which, int32_t* region, int32_t* group, int64_t* ncells,
                       int64_t* nfaces, int32_t* nbound);
/* boundary group idx of that view, in file order: name (NUL-terminated, name_cap bytes
 * available), first face and face count */
int  hs_mesh_proc_boundary(const hs_mesh* m, int32_t which, int32_t idx, char* name, int32_t name_cap,
                           int64_t* start, int64_t* nfaces);
void hs_mesh_free(hs_mesh* m);

#ifdef __cplusplus
}
#endif
#endif
