/*
 * nuif_io.h - frame ingest: the reference's on-disk RGB-D frame format and a
 * prefetching loader that decodes recorded frames into pinned host memory and
 * uploads them to the device ahead of the consumer (libnuif_kinfu.so).
 *
 * File format (src/Foundation/NuiFileIOUtilities.cpp:13-110), little endian:
 *   int64 timestamp, uint32 width, uint32 height, uint32 stored size, payload.
 *   The payload is the raw pixel buffer when the stored size equals
 *   width * height * bytes per pixel, else the zlib stream of it (compress2 at
 *   Z_BEST_SPEED).  A recording is a directory of <id>.depth (uint16 mm),
 *   <id>.color (BGRA), <id>.cmap (float2 colour-space points), <id>.vertex
 *   (float3 camera-space points) and <id>.camIntri (float32 fx, fy, cx, cy)
 *   for id = 0, 1, 2, ... (src/Frame/NuiFrameSaveManager.cpp:24-80); it ends at
 *   the first id for which no file exists
 *   (src/DeviceManager/FileLoader/NuiFrameLoader.cpp:26-91).
 *
 * Every function returns 0 (or a documented non-negative value) on success and
 * a negative status on failure; nuifk_last_error() gives the calling thread's
 * message.  Unlike the reference, a truncated file, a size mismatch or a zlib
 * stream that does not inflate to exactly the expected size is an error
 * (NUIF_ERR_IO) rather than a buffer left partly unwritten.
 */
#ifndef NUIF_IO_H
#define NUIF_IO_H

#include <stddef.h>
#include <stdint.h>

#include "nuif.h"

#ifdef __cplusplus
extern "C" {
#endif

#define NUIF_ERR_IO -6          /* file missing, truncated or not decodable */

/* parts of a recorded frame (bits of nuif_loaded_frame.parts) */
#define NUIF_FRAME_DEPTH 1      /* <id>.depth   uint16 per pixel */
#define NUIF_FRAME_COLOR 2      /* <id>.color   BGRA per pixel */
#define NUIF_FRAME_CMAP 4       /* <id>.cmap    2 x float32 per pixel */
#define NUIF_FRAME_VERTEX 8     /* <id>.vertex  3 x float32 per pixel */
#define NUIF_FRAME_CAMERA 16    /* <id>.camIntri */

/* Foundation/NuiFileIOUtilities.cpp:13-52 writeFrameImage: header, then the raw buffer, or
 * (compressed != 0) its zlib stream when compress2 succeeds. */
int nuif_io_write_frame_image(const char* path, int compressed, int64_t timestamp, uint32_t width,
                              uint32_t height, const void* buffer, uint32_t bytes);
/* :54-75 readFrameImageHeader (each output may be NULL) */
int nuif_io_read_frame_header(const char* path, int64_t* timestamp, uint32_t* width, uint32_t* height);
/* :77-110 readFrameImageBuffer: `bytes` is the raw size the caller expects
 * (width * height * bytes per pixel). */
int nuif_io_read_frame_buffer(const char* path, void* buffer, uint32_t bytes);
/* :112-160 writeCamera / readCamera: 4 float32 */
int nuif_io_write_camera(const char* path, float fx, float fy, float cx, float cy);
int nuif_io_read_camera(const char* path, float* fx, float* fy, float* cx, float* cy);

/* One recorded frame in the loader's memory (valid until nuif_frame_loader_release). */
typedef struct nuif_loaded_frame {
    uint32_t frame_id;
    uint32_t parts;                 /* NUIF_FRAME_* bits of the files that loaded */
    int32_t slot;                   /* pass to nuif_frame_loader_release */
    uint32_t pad;
    int64_t timestamp[4];           /* depth, color, cmap, vertex */
    uint32_t width[4], height[4];   /* depth, color, cmap, vertex */
    const uint16_t* depth;          /* host: pinned when the loader uploads */
    const uint8_t* color;
    const float* cmap;
    const float* vertex;
    const uint16_t* device_depth;   /* device copies (NULL unless the loader uploads) */
    const uint8_t* device_color;
    float fx, fy, cx, cy;           /* .camIntri (0 when absent) */
} nuif_loaded_frame;

/* DeviceManager/FileLoader/NuiFrameLoader.cpp:26-91 as a prefetching pipeline: `threads`
 * workers read and inflate frames id, id+1, ... into a ring of `slots` buffers, ahead of the
 * consumer.  upload != 0: host buffers are pinned, and each worker copies the depth and colour
 * to device buffers of its slot on its own stream and waits for the copy before it hands the
 * frame over, so a frame returned by next is already resident in HBM (nuif_init first).
 * upload == 0 needs no GPU. */
int nuif_frame_loader_create(const char* dir, int slots, int threads, int upload, void** out);
/* The next frame in id order: 1 and *out filled, 0 at the end of the recording, < 0 on a
 * read error (the frame's id and file are in the message).  Blocks until it is decoded. */
int nuif_frame_loader_next(void* loader, nuif_loaded_frame* out);
/* Hands the slot back for prefetching. */
int nuif_frame_loader_release(void* loader, int slot);
/* NuiFrameLoader::reset: back to frame 0 (outstanding frames become invalid). */
int nuif_frame_loader_reset(void* loader);
int nuif_frame_loader_destroy(void* loader);

/* Frame/NuiFrameSaveManager.cpp:24-80 process(): writes the parts of `frame` listed in its
 * `parts` as <dir>/<id>.* with id = frame->frame_id (host pointers are read). */
int nuif_frame_save(const char* dir, const nuif_loaded_frame* frame, int compressed);

#ifdef __cplusplus
}
#endif

#endif /* NUIF_IO_H */
