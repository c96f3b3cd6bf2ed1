from .primitives import (Primitive, Primitives, Sphere, Capsule, RollingPin, RollingPinExt, Chopsticks, Box,
                         Gripper, Gripper2, Prism, Knife, Cylinder, Torus)

__all__ = ["Primitive", "Primitives", "Sphere", "Capsule", "RollingPin", "RollingPinExt", "Chopsticks", "Box",
           "Gripper", "Gripper2", "Prism", "Knife", "Cylinder", "Torus"]
