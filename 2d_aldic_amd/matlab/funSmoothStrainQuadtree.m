function F = funSmoothStrainQuadtree(F,DICmesh,DICpara)
% Library-backed replacement of funSmoothStrainQuadtree.m:1-104 (main_ALDIC_Quadtree.m:172,
% 221, 283): two passes of regularizeNd of each F component over the grid
% min:elementMinSize:max, the 3x3 Gaussian nanconv ('edge') and the read-back at the
% nodes, on the GPU.  Non-interactive: an empty StrainFilterStd or StrainFilterSize takes
% the reference's default (std 0.5, size 2*ceil(2*std)+1) instead of prompting.
try smoothness = DICpara.StrainSmoothness; catch, smoothness = 1e-5; end
try fstd = DICpara.StrainFilterStd; catch, fstd = []; end
try fsize = DICpara.StrainFilterSize; catch, fsize = []; end
if isempty(fstd), fstd = 0; end
if isempty(fsize), fsize = 0; end
F = aldic_mex('smooth_strain_quadtree', full(F), DICmesh.coordinatesFEM, ...
    DICmesh.elementMinSize, smoothness, fstd, fsize);
end
