/*
 * ugmg_mgio.h -- UG's multigrid files ("<name>.ug.mg.<xdr|asc|bin>", written
 * by SaveMultiGrid, gm/ugio.cc:1724, read by LoadMultiGrid, gm/gm.h:3466),
 * host only.  Sequential files (nparfiles 1, MGIO_DEBUG 0), versions
 * UG_IO_2.2 .. 2.5, 2D and 3D, in the three encodings of low/bio.cc.
 *
 * Sections (gm/mgio.cc): the head (title and mode in ASCII; version,
 * ident, domain, multigrid and format names, 11 counts), the element types
 * (corners, edges and sides of each tag), the refinement rules the file
 * uses, the coarse grid (points, then elements with corner node IDs), the
 * domain's boundary points (a section of counted length, skipped: the
 * coarse points already hold every position), then per coarse element with
 * refinements its refinement records, depth first over the sons
 * (SetHierRefinement / InsertLocalTree, gm/ugio.cc:1070-1100, 2299-2625).
 *
 * The multigrid is rebuilt as LoadMultiGrid's InsertLocalTree does: each
 * record names its rule and the node IDs of the element's new corners --
 * the sons of its corners, then the midnodes of the edges the rule refines,
 * the side nodes (3D), the centre node -- and the sons take their corners
 * from that list.  A node ID met again is the same node (a neighbour made
 * it).  Positions: a corner's son sits on its father, a midnode at its
 * edge's midpoint, a side node at its side's centre, a centre node at the
 * element's centre (the images of those local points under the element
 * map for straight-sided elements), a moved corner where the record puts
 * it.  New boundary nodes of a curved domain would need the domain's
 * boundary parametrisation, which the file does not hold.
 *
 * Per level the result lists the nodes in node-ID order -- the order UG's
 * RenumberNodes (gm/ugio.cc:199-250) numbered them on the grid that wrote
 * the file: level 0 by vertex ID (boundary vertices first), finer levels
 * in their list order -- the elements in creation order, and the
 * interpolation rows StandardIntCorNodeVector (np/algebra/transgrid.cc:
 * 235-336) forms for node vectors: weight 1 from the father node for a
 * corner's son, else GNs at the node's local coordinate in the element that
 * created it, over that element's corners in corner order, zeros skipped.
 */
#ifndef UGMG_MGIO_H
#define UGMG_MGIO_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ugmg_mg ugmg_mg;

/* General information of a read file (MGIO_MG_GENERAL, gm/mgio.h). */
typedef struct {
  int mode;            /* 0 XDR, 1 ASCII, 2 BIN */
  char version[128];
  char ident[4096];
  char domain[128], multigrid[128], format[128];
  int dim, magic_cookie, heapsize;
  int nlevels, nnodes, npoints, nelements;   /* as the file states them */
  int nrules, ncoarse_points, ncoarse_bnd_points, ncoarse_elements;
} ugmg_mg_info;

/* Read and rebuild a multigrid file.  0, or -1 with a message in err (up
 * to errlen bytes) when the file does not follow the layout or uses what
 * is not read (parallel files, debug extensions, element types whose
 * transfer weights are not formed), -2 when it cannot be opened. */
int ugmg_mg_read(const char *path, ugmg_mg **out, char *err, int errlen);
void ugmg_mg_free(ugmg_mg *mg);
int ugmg_mg_get_info(const ugmg_mg *mg, ugmg_mg_info *info);

/* Level `level` of the rebuilt multigrid.  Any output may be NULL; sizes
 * first: *nnodes, *nelem, *ncorners (sum over the elements), *np (entries
 * of the interpolation rows; 0 on level 0).
 *   ids[nnodes]          node IDs, ascending (the level's vector order)
 *   xyz[nnodes*dim]      positions
 *   elem_tag[nelem]      UG element tags (TRIANGLE 3, QUADRILATERAL 4 in 2D;
 *                        TETRAHEDRON 4, PYRAMID 5, PRISM 6, HEXAHEDRON 7)
 *   elem_ptr[nelem+1], elem_corners[ncorners]   corners as level indices
 *   p_ptr[nnodes+1], p_idx[np], p_w[np]          interpolation from level-1
 *                                                (columns: level-1 indices)
 */
int ugmg_mg_level(const ugmg_mg *mg, int level, int *nnodes, int *nelem, int *ncorners, int *np,
                  int *ids, double *xyz, int *elem_tag, int *elem_ptr, int *elem_corners,
                  int *p_ptr, int *p_idx, double *p_w);

#ifdef __cplusplus
}
#endif

#endif
