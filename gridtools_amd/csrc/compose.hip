// Stencil compositions: several multistages of one run() on one stream.
//
// hdiff -> vadv: the horizontal diffusion of a field feeds the vertical advection as its u_stage
// (u_stage = hdiff(in, coeff); then vadv(utens_stage, u_stage, wcon, u_pos, utens)).  In the reference this is
// one spec with multi_pass(hdiff MSS, forward MSS, backward MSS); its gpu backend launches the parallel hdiff
// multistage and the fused vadv multistages separately (stencil/gpu/entry_point.hpp:96-124: a multistage whose
// inputs are read with a horizontal offset or a k offset cannot join the previous launch), and so does this one.
#include <hip/hip_runtime.h>

#include "common.hpp"

extern "C" int gt355_hdiff_vadv(int precision, const gt355_field *in, const gt355_field *coeff,
    const gt355_field *u_stage, const gt355_field *utens_stage, const gt355_field *wcon, const gt355_field *u_pos,
    const gt355_field *utens, double dtr_stage, int ni, int nj, int nk, void *stream) {
    GT355_CHECK(in && coeff && u_stage && utens_stage && wcon && u_pos && utens, "gt355_hdiff_vadv: null field");
    GT355_CHECK(u_stage->origin != in->origin, "gt355_hdiff_vadv: u_stage must not alias in");
    // validate the second multistage's fields before the first one launches, so a bad call changes nothing
    GT355_CHECK(nk >= 2, "gt355_hdiff_vadv: nk = %d, the vertical sweep needs at least 2 levels", nk);
    const int dt = in->dtype;
    int rc;
    if ((rc = gt355::check_field(utens_stage, "utens_stage", dt, 0, ni, 0, nj, 0, nk)) ||
        (rc = gt355::check_field(wcon, "wcon", dt, 0, ni + 1, 0, nj, 0, nk)) ||
        (rc = gt355::check_field(u_pos, "u_pos", dt, 0, ni, 0, nj, 0, nk)) ||
        (rc = gt355::check_field(utens, "utens", dt, 0, ni, 0, nj, 0, nk)))
        return rc;
    rc = gt355_hdiff(precision, u_stage, in, coeff, ni, nj, nk, stream);
    if (rc)
        return rc;
    return gt355_vadv(precision, utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, ni, nj, nk, stream);
}
