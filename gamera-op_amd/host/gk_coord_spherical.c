/*
 * Spherical coordinates for a program linked against the host facade.
 *
 * The reference picks its geometry when it is built: COORD_TYPE=2 compiles
 * src/spherical/*.c in place of src/cartesian/*.c (CMakeLists.txt:77-90).
 * The facade holds both geometries and uses Cartesian metrics unless this
 * file is compiled into the program, as a spherical problem's build adds it
 * beside the problem file; it then builds spherical metrics, weights and
 * sources (src/spherical/spherical_general.c) for every grid.
 */
extern int gk_coord_type;

__attribute__((constructor)) static void gk_select_spherical(void) { gk_coord_type = 2; }
