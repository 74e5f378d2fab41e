"""`pycujpegxl`: the reference's Python module name over this library.

Same names, signatures and errors as python/pycujpegxl.cc:81-90 of the
reference (`API_VERSION`, `api_version`, `query_backend`, `encode`,
`encode_with_stats`): callers such as its encode_cli.py and its ladder
conformance test run unchanged.  Encoding goes through the C ABI of
libcujpegxl.so.1 (host NV12 via cujpegxl_encoder_encode_host).
"""

from cujpegxl_amd import API_VERSION, api_version, encode, encode_with_stats, query_backend

__all__ = ["API_VERSION", "api_version", "query_backend", "encode", "encode_with_stats"]
