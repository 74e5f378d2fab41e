CUJPEGXL_STAGES_1 {
    global:
        cujpegxl_*;
        cjx_*;
    local:
        *;
};
