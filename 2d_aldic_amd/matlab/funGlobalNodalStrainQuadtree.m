function [StrainNodalPt,StrainGaussPt,CoordGaussPt] = funGlobalNodalStrainQuadtree(...
    DICmesh,U,GaussPtOrder,waitBarDisplayOrNot)
% Library-backed replacement of funGlobalNodalStrainQuadtree.m:31-255 (smoothness 0): the
% 2x2 Gauss-point strains of the quadtree mesh (midside nodes dropped) on the GPU, then
% MATLAB's own scatteredInterpolant(...,'linear') from the Gauss points to the nodes, whose
% barycentric weights are handed to the library (the reference's interpolation exactly).
d = [];  % Subpb2Quadtree.m:32: DICmesh.dirichlet (= markCoordHoleEdge, GenerateQuadtreeMesh.m:116)
if isfield(DICmesh, 'dirichlet'), d = DICmesh.dirichlet(:); end
aldic_mex('set_quadtree', DICmesh.coordinatesFEM, DICmesh.elementsFEM, d);
E = size(DICmesh.elementsFEM, 1);
[~, StrainGaussPt, CoordGaussPt] = aldic_mex('nodal_strain_quadtree', U(:), E);
StrainNodalPt = zeros(4 * size(DICmesh.coordinatesFEM, 1), 1);
for k = 1:4
    Fi = scatteredInterpolant(CoordGaussPt(:, 1), CoordGaussPt(:, 2), StrainGaussPt(:, k), 'linear');
    StrainNodalPt(k:4:end) = Fi(DICmesh.coordinatesFEM(:, 1), DICmesh.coordinatesFEM(:, 2));
end
end
