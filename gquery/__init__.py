"""gQuery's module paths over the MI355X engine (gquery_amd): a program written against the
reference imports gquery.shapes.mesh.Mesh, gquery.shapes.polyline.Polyline, the records of
gquery.shapes.primitive, the array types of gquery.core.fwd and the builders of gquery.gquery_ext
unchanged, and runs the HIP kernels."""
