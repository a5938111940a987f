"""Sphere collision constraint (plainmp.constraint; constraints/primitive_sphere_collision.hpp).

``SphereCollisionCst`` keeps the reference's single-configuration methods
(``is_valid``, ``evaluate``, ``update_kintree``, ``get_all_spheres``,
``get_group_spheres``) and adds batched ones (``is_valid_batch``,
``evaluate_batch``, ``sphere_clearance_batch``, ``sphere_centers_batch``,
``pair_clearance_batch``).
"""

from ._native.constraint import SphereAttachmentSpec, SphereCollisionCst  # noqa: F401

__all__ = ["SphereAttachmentSpec", "SphereCollisionCst"]
