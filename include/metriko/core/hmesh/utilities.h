// metriko (MI355X build) -- uv <-> surface helpers on an Hmesh.
//
// Drop-in for conversion_2d_3d of the reference's core/hmesh/utilities.h: the surface
// point of a face at a uv, by the face's corner uv (corner of face.half() first, then of
// its next and prev halfedges) and the corners' vertex positions.
//
// Restated from komietty/metriko include/metriko/core/hmesh/utilities.h
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_HMESH_UTILITIES_H
#define METRIKO_HMESH_UTILITIES_H
#include "../common/utilities.h"
#include "hmesh.h"

namespace metriko {
    inline Row3d conversion_2d_3d(const complex origin2d, const complex p1_2d, const complex p2_2d,
                                  const Row3d& origin3d, const Row3d& p1_3d, const Row3d& p2_3d, const complex uv) {
        const complex c = calc_coefficient(origin2d, p1_2d, p2_2d, uv);
        return origin3d + (p1_3d - origin3d) * c.real() + (p2_3d - origin3d) * c.imag();
    }
    inline Row3d conversion_2d_3d(const Face& face, const VecXc& cfn, const complex uv) {
        const Crnr c1 = face.half().crnr(), c2 = face.half().next().crnr(), c3 = face.half().prev().crnr();
        return conversion_2d_3d(cfn(c1.id), cfn(c2.id), cfn(c3.id), c1.vert().pos(), c2.vert().pos(),
                                c3.vert().pos(), uv);
    }
}
#endif
