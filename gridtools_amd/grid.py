"""halo_descriptor and make_grid, as in the reference (common/halo_descriptor.hpp:44-227,
stencil/frontend/make_grid.hpp:19-42, stencil/core/grid.hpp:30-128)."""
from dataclasses import dataclass


@dataclass(frozen=True)
class halo_descriptor:
    """(minus, plus, begin, end, total_length); end is inclusive (common/halo_descriptor.hpp:44-49)."""

    minus: int
    plus: int
    begin: int
    end: int
    total_length: int

    def __post_init__(self):
        # the constructor's assertions (common/halo_descriptor.hpp:88-93)
        if not (self.minus + self.plus + (self.end - self.begin + 1) <= self.total_length):
            raise ValueError(f"halo_descriptor {self}: minus + plus + size exceeds total_length")
        if not self.begin <= self.end:
            raise ValueError(f"halo_descriptor {self}: begin > end")
        if not self.plus <= self.total_length - self.end - 1:
            raise ValueError(f"halo_descriptor {self}: plus halo exceeds total_length")
        if not self.begin >= self.minus:
            raise ValueError(f"halo_descriptor {self}: begin < minus")

    @classmethod
    def of_size(cls, size):
        """halo_descriptor(size): no halo."""
        return cls(0, 0, 0, size - 1, size)

    @classmethod
    def symmetric(cls, halo, total):
        """The descriptor the reference tests build: (h, h, h, total - h - 1, total) (test_environment.hpp:198)."""
        return cls(halo, halo, halo, total - halo - 1, total)

    def as_tuple(self):
        return (self.minus, self.plus, self.begin, self.end, self.total_length)


@dataclass(frozen=True)
class Grid:
    i_start: int
    i_size: int
    j_start: int
    j_size: int
    k_size: int

    @property
    def origin(self):
        return (self.i_start, self.j_start, 0)


def make_grid(hi, hj, nk):
    """make_grid(halo_descriptor, halo_descriptor, k size) (frontend/make_grid.hpp:19-42): the compute domain
    is [begin, end] in i and j and [0, nk) in k."""
    if isinstance(hi, int):
        hi = halo_descriptor.of_size(hi)
    if isinstance(hj, int):
        hj = halo_descriptor.of_size(hj)
    return Grid(hi.begin, hi.end - hi.begin + 1, hj.begin, hj.end - hj.begin + 1, int(nk))
