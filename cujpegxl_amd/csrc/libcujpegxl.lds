CUJPEGXL_1 {
    global:
        cujpegxl_api_version;
        cujpegxl_status_string;
        cujpegxl_backend_string;
        cujpegxl_query_backend;
        cujpegxl_max_output_size;
        cujpegxl_encoder_create;
        cujpegxl_encoder_configure;
        cujpegxl_encoder_encode;
        cujpegxl_encoder_destroy;
    local:
        *;
};

CUJPEGXL_2 {
    global:
        cujpegxl_ext_version;
        cujpegxl_encoder_encode_host;
        cujpegxl_encoder_encode_async;
        cujpegxl_encoder_finish;
} CUJPEGXL_1;

CUJPEGXL_3 {
    global:
        cujpegxl_noalloc_workspace_bytes;
        cujpegxl_encode_noalloc;
} CUJPEGXL_2;

CUJPEGXL_4 {
    global:
        cujpegxl_encoder_set_mode;
} CUJPEGXL_3;
