CUJPEGXL_1 {
    global:
        cujpegxl_*;
    local:
        *;
};
