// Jacobian (nodal and raster basis) and Raster_Pixel entry points.
#include "solver.hpp"

extern "C" {
int tmi_jacobian_nodal(tmi_mesh, int, int, const double *, const double *, const double *, int, double *)
{
    return TMI_ERR_UNSUPPORTED;
}
int tmi_raster_create(tmi_mesh, const int *, tmi_raster *) { return TMI_ERR_UNSUPPORTED; }
int tmi_raster_free(tmi_raster) { return TMI_OK; }
int tmi_raster_dims(tmi_raster, int *, int *) { return TMI_ERR_UNSUPPORTED; }
int tmi_raster_get_elref(tmi_raster, int *, int *) { return TMI_ERR_UNSUPPORTED; }
int tmi_raster_map_mesh_to_grid(tmi_raster, const double *, double *) { return TMI_ERR_UNSUPPORTED; }
int tmi_jacobian_raster(tmi_raster, int, int, const double *, const double *, const double *, int, double *)
{
    return TMI_ERR_UNSUPPORTED;
}
}
