"""The parts of the reference's ``montblanc.util`` on the path's input side
(montblanc/util/__init__.py:28, 340-342): the sky-model file reader and
``shape_list``."""

from .sky_model import ParseResults, parse_sky_model, shape_list, source_arrays  # noqa: F401
