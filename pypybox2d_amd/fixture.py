"""Fixture: a shape attached to a body plus its material and filter data.

Same constructor and copy semantics as pypybox2d/fixture.py:55-85 (the shape
is copied on construction; copying a fixture copies it again).  Sensors are
accepted here but rejected by the device path (b2g_upload reports
B2G_E_UNSUPPORTED): the contact pipeline does not evaluate sensor overlaps.
"""
from __future__ import annotations

from copy import copy

from .common import EmptyFixtureError, is_valid_float

__all__ = ("Fixture",)


class Fixture(object):
    def __init__(self, shape=None, friction=0.2, restitution=0.0, density=0.0, category_bits=0x0001,
                 mask_bits=0xFFFF, group_index=0, sensor=False, user_data=None, body=None):
        self._friction = friction
        self._restitution = restitution
        self._density = density
        self._sensor = sensor
        self._category_bits = category_bits
        self._mask_bits = mask_bits
        self._group_index = group_index
        self._shape = copy(shape)
        self.user_data = user_data
        if body is not None and self._shape is None:
            raise EmptyFixtureError("Cannot attach empty fixture to a body")
        self._body = body

    def __copy__(self):
        return Fixture(shape=self._shape, friction=self._friction, restitution=self._restitution,
                       density=self._density, category_bits=self._category_bits, mask_bits=self._mask_bits,
                       group_index=self._group_index, sensor=self._sensor, user_data=self.user_data)

    def __repr__(self):
        return "Fixture(shape=%s, friction=%g, restitution=%g, density=%g)" % (
            self._shape, self._friction, self._restitution, self._density)

    @property
    def body(self):
        return self._body

    @property
    def shape(self):
        return self._shape

    @property
    def mass_data(self):
        return self._shape.compute_mass(self._density)

    def _changed(self):
        if self._body is not None and self._body._world is not None:
            self._body._world._topology_changed()

    @property
    def friction(self):
        return self._friction

    @friction.setter
    def friction(self, value):
        # like the reference, existing contacts keep their mixed friction
        self._friction = value
        self._changed()

    @property
    def restitution(self):
        return self._restitution

    @restitution.setter
    def restitution(self, value):
        self._restitution = value
        self._changed()

    @property
    def density(self):
        return self._density

    @density.setter
    def density(self, density):
        if not is_valid_float(density) or density < 0:
            raise ValueError("Invalid density")
        self._density = density

    @property
    def sensor(self):
        return self._sensor

    @sensor.setter
    def sensor(self, sensor):
        # fixture.py:169-174: wakes the body; the next step rejects sensors
        if sensor != self._sensor:
            if self._body is not None:
                self._body.awake = True
            self._sensor = sensor
            self._changed()

    @property
    def category_bits(self):
        return self._category_bits

    @category_bits.setter
    def category_bits(self, category_bits):
        self._category_bits = category_bits
        self._refilter()

    @property
    def mask_bits(self):
        return self._mask_bits

    @mask_bits.setter
    def mask_bits(self, mask_bits):
        self._mask_bits = mask_bits
        self._refilter()

    @property
    def group_index(self):
        return self._group_index

    @group_index.setter
    def group_index(self, group_index):
        self._group_index = group_index
        self._refilter()

    def _refilter(self):
        """fixture.py:269-285: the fixture's contacts are tested against the filter again at the next
        step where one of their bodies is awake, and the pair search reconsiders its pairs."""
        world = self.world
        if world is not None:
            if world._locked:  # inside a listener callback: it takes effect from the next step on
                world._deferred_refilter.append(self)
            else:
                world._refilter(self)

    def test_point(self, p):
        return self._shape.test_point(self._body._xf, p)

    @property
    def world(self):
        return self._body._world if self._body is not None else None
