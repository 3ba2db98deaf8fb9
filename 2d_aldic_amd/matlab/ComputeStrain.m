% ==================================
% Library-backed replacement of the ComputeStrain.m script (ComputeStrain.m:11-139),
% run from main_ALDIC.m's Section 8 with the same workspace variables: DICpara
% (MethodToComputeStrain, StrainType, winstepsize), ULocal, FLocal, USubpb2,
% coordinatesFEM, elementsFEM, M, N.  Sets FStrainFinite, FStraintemp, FStrainWorld,
% Rad and temp3 as the script does (FStrain is FStrainFinite before the strain type, so
% only for StrainType 0 is it the same vector).
% ==================================
switch DICpara.MethodToComputeStrain
    case 2
        prompt = 'What is your half window size: ';
        Rad = input(prompt);
    case 3
        aldic_mex('set_mesh', coordinatesFEM, elementsFEM(:, 1:4));
        Rad = 1;
    case 0
        Rad = 0;
    otherwise
        Rad = 1;
end
if DICpara.MethodToComputeStrain == 3, UStrainIn = USubpb2; else, UStrainIn = ULocal; end
if DICpara.MethodToComputeStrain == 0, FStrainIn = FLocal; else, FStrainIn = []; end
[FStrainFinite, FStrainWorld] = aldic_mex('compute_strain', M, N, DICpara.winstepsize, ...
    UStrainIn(:), FStrainIn(:), DICpara.MethodToComputeStrain, Rad, DICpara.StrainType);
if DICpara.StrainType == 0, FStrain = FStrainFinite; end
temp = reshape((1:M*N)', M, N); temp2 = temp(Rad+1:M-Rad, Rad+1:N-Rad); temp2 = temp2(:);
temp3 = reshape([4*temp2-3, 4*temp2-2, 4*temp2-1, 4*temp2]', [], 1);
FStraintemp = FStrainWorld; FStraintemp(2:4:end) = -FStraintemp(2:4:end);
FStraintemp(3:4:end) = -FStraintemp(3:4:end);
