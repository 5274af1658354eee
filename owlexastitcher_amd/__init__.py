"""MI355X-native exaStitch volume path tracer (hot path of owlExaStitcher).

All computation lives in the in-tree HIP library ``libexastitch.so``
(``owlexastitcher_amd/csrc``, C ABI in ``include/exa/exa.h``); this package
is its Python host mirror.  Importing the host classes loads the library and
raises ImportError if it has not been built.
"""
from ._capi import LIB_PATH, ExaError, lib  # noqa: F401
from .host import (ACCEL_BVH, ACCEL_TABLES, DIRECT_LIGHTING, PATH_TRACER, RAY_MARCHER, RENDER_MEGAKERNEL,  # noqa: F401
                   ELEMENT_IN_WAVE, ELEMENT_QUEUE_ROUNDS, ELEMENT_WORKGROUP_QUEUE, RENDER_PERSISTENT, AmrScene, DualMesh, ExaBrickModel, ExaBrickSampler, ExaStitchModel,
                   ExaStitchSampler, NeuralField, OWLRenderer,
                   RenderBatch,
                   camera_look_at, make_grids,
                   tf_rasterize, xf_load, xf_save)
