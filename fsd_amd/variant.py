"""Box-deformation variant of the reference (variant.py, VariantShearFunction.cc):
the strain of a shear protocol wrapped into [-max_strain, max_strain), held at
0 before the protocol's offset and at its final value after total_timestep.
System.box_resize(xy=shear_variant(...)) turns it into the box tilt factor
each timestep, as HOOMD's update.box_resize does with this variant."""
import math


class shear_variant:
    def __init__(self, function_form, total_timestep, max_strain=0.5):
        if total_timestep <= 0:
            raise RuntimeError("Error creating variant: total_timestep must be positive")
        self.function_form = function_form
        self.total = int(total_timestep)
        self.min_value, self.max_value = -float(max_strain), float(max_strain)
        self.range = self.max_value - self.min_value
        self.offset = function_form.get_offset()
        self.end_value = self._wrap(function_form.get_strain(self.offset + self.total))

    def _wrap(self, v):
        return v - self.range * math.floor((v - self.min_value) / self.range)

    def get_value(self, timestep):
        if timestep < self.offset:
            return 0.0
        if timestep >= self.offset + self.total:
            return self.end_value
        return self._wrap(self.function_form.get_strain(timestep))

    __call__ = get_value
