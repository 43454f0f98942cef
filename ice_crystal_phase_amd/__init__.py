"""MI355X-native polarized geometric-optics phase-function generator.

Hot path: trace rays through water drops and hexagonal ice prisms, split them
with FP32/FP64 Fresnel-Jones algebra, bin the outgoing flux into fixed-point
(u, phi) histograms and build FP64 CDF records, all in hand-written gfx950
kernels behind the C-ABI in include/icephase.h (libicephase.so).
"""
from .icephase import (AtlasMass, AuditError, Batch, CdfPolicy, CdfReport, Coherent, Context, HistStats, IcePhaseError, Record,
                       REC_DIFFRACTION, RUN_CDF, RUN_GRAPH, RUN_NO_AUDIT, RUN_TRACE, SHAPE_DROP, SHAPE_HEX,
                       SHAPE_SPHERE, STATUS_NAMES, drop_incident, drop_projected_area, drop_record, drop_shape,
                       hex_projected_area, hex_record, lib, quaternion_rotation, record_rays, sphere_record,
                       u_edges, water_index, water_optics, write_record)

__all__ = ["AtlasMass", "AuditError", "Batch", "CdfPolicy", "CdfReport", "Coherent", "Context", "HistStats", "IcePhaseError", "Record",
           "REC_DIFFRACTION", "RUN_CDF", "RUN_GRAPH", "RUN_NO_AUDIT", "RUN_TRACE", "SHAPE_DROP", "SHAPE_HEX",
           "SHAPE_SPHERE", "STATUS_NAMES", "drop_incident", "drop_projected_area", "drop_record", "drop_shape",
           "hex_projected_area", "hex_record", "lib", "quaternion_rotation", "record_rays", "sphere_record",
           "u_edges", "water_index", "water_optics", "write_record"]
