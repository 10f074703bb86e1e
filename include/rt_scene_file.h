/*
 * rt_scene_file.h — the reference's scene-description front end, in the
 * same library (librt_hip.so) and without Boost, Qt or OpenCV.
 *
 * The reference reads a scene in three steps (src/main.cpp:150-182):
 *   SceneGrammar::parse   the text grammar (src/sceneparser.cpp:37-125,
 *                         src/sceneparser.cpp:129-156)
 *   resolveVars           obj(...) replaced by the OBJ file's triangles, its
 *                         MTL materials prepended as assignments
 *                         (src/sceneast.cpp:731-747, src/objreader.cpp:142-366)
 *   buildScene            assignments evaluated in order, root list = lights
 *                         then "intersectable" (src/sceneast.cpp:589-695)
 * and, for --dump-bvh, createBVH + SceneDumper (src/sceneast.cpp:749-766,
 * src/bvh.cpp:97-133, src/scenedumper.cpp).
 *
 * rt_scene_file_load / _parse do the first two steps; rt_scene_file_scene
 * does the third and hands the result over as the rt_scene_desc that
 * rt_scene_upload takes (include/rt_scene.h), with the camera variable.
 * rt_scene_file_dump_bvh writes the --dump-bvh text.
 *
 * Planes, boxes and CSG trees become the analytic objects of rt_scene_desc,
 * texture materials its textures (PNG, PPM or BMP images).  What the GPU path
 * does not render (spheres under non-similarity instance transforms, boxes
 * under instance transforms other than axis scale and translation) still
 * parses and dumps;
 * rt_scene_file_scene then returns RT_ERR_UNSUPPORTED naming the
 * construct.  Input the reference accepts
 * with a warning on stderr (unknown material variable -> DarkMatter,
 * unreadable OBJ -> empty list, unreadable or undecodable texture -> DarkMatter,
 * unreadable environment map -> BlackLight) is accepted the same way, the
 * warning going to rt_scene_file_messages.
 */
#ifndef RT_SCENE_FILE_H
#define RT_SCENE_FILE_H

#include <stdint.h>
#include "rt_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rt_scene_file rt_scene_file;

/* Parse the scene file at `path`; obj(...) paths are relative to its
 * directory (src/sceneast.cpp:798-826).  RT_ERR_IO: unreadable file;
 * RT_ERR_PARSE: text the grammar rejects, or an OBJ face index out of
 * range (the reference reads out of bounds there); details in
 * rt_scene_file_last_error(). */
int rt_scene_file_load(const char* path, rt_scene_file** out);

/* The same from a string; obj(...) paths are relative to base_dir (NULL =
 * the current directory). */
int rt_scene_file_parse(const char* text, const char* base_dir, rt_scene_file** out);

/* Message of the last failed load/parse on this thread. */
const char* rt_scene_file_last_error(void);

/* Warnings collected while loading and evaluating (one per line). */
const char* rt_scene_file_messages(const rt_scene_file* f);

/* Evaluate the assignments (src/sceneast.cpp:673-695) into a scene
 * description and the "camera" variable.  The arrays `desc` points to are
 * owned by `f` and stay valid until rt_scene_file_free.  Triangles come as
 * a triangle soup (three vertices each); a quad is two triangles with one
 * primitive ID; an instance is one primitive ID.  Errors: RT_ERR_INVALID
 * when "camera" or "intersectable" is not defined, RT_ERR_UNSUPPORTED for
 * constructs the GPU path lacks (rt_scene_file_last_error names it). */
int rt_scene_file_scene(rt_scene_file* f, rt_scene_desc* desc, rt_camera* camera);

/* --dump-bvh: every intersectable variable flattened and rebuilt as b(l, r,
 * box(...)) nodes by the reference's median split, all assignments printed
 * in the reference's format.  Writes to `path`. */
int rt_scene_file_dump_bvh(const rt_scene_file* f, const char* path);

/* The same text into buf (cap bytes, NUL-terminated when it fits); returns
 * the text length, or a negative rt_status. */
int64_t rt_scene_file_dump_bvh_text(const rt_scene_file* f, char* buf, int64_t cap);

void rt_scene_file_free(rt_scene_file* f);

#ifdef __cplusplus
}
#endif

#endif /* RT_SCENE_FILE_H */
