"""Geometry: shapes (src/prismo/geometry/__init__.py) and their device
rasterisation into update coefficients."""

from .shapes import Box, CustomShape, Cylinder, GeometryGroup, Material, Polygon, Shape, Sphere

__all__ = ["Shape", "Material", "Box", "Sphere", "Cylinder", "Polygon", "CustomShape",
           "GeometryGroup"]
