"""gquery.shapes.primitive (gquery/shapes/primitive.py:9-84): the query records."""
from gquery_amd.shapes.primitive import (  # noqa: F401
    BoundarySamplingRecord, BoundarySamplingRecord3D, ClosestPointRecord, ClosestPointRecord3D,
    ClosestSilhouettePointRecord, ClosestSilhouettePointRecord3D, Intersection, Intersection3D)
